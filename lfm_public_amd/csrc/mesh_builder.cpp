// mesh_builder.cpp -- host set-up of one partition (polyMesh -> SoA layout).
//
// Restates, in a face-list form, what the reference spreads over:
//   polyMeshReaderOF/polyMeshReaderOF.cpp:47-59     empty-face masking
//   polyMeshReaderOF/polyMeshReaderOF.cpp:149-162   face neighbour (cyclic twins)
//   polyMeshReaderOF/polyMeshReaderOF.cpp:226-248   cyclic twin index
//   polyMeshReaderOF/polyMeshReaderOF.cpp:501-552   boundary/interior submesh split
//   src/mesh_reader.cpp:163-260                     per-cell valid faces, neighbour typing
//   src/mesh_reader.cpp:264-338                     face area vectors, owner sign
//   src/mesh_reader.cpp:375-456                     cell -> neighbour vectors (cyclic, mirror)
//   src/cfd_v0.cpp:346-431                          boundary ghost cells
//   src/cfd_v0.cpp:760-861                          S, d, linear weight, 1/V
//   src/cfd_v0.cpp:278-339                          valid-first face order
//   src/cfd_v0.cpp:867-967                          sponge sigma, initial conservatives
//   src/cfd_v0.cpp:432-540, 587-700                 MPI ghosts, send lists
//   src/mesh_reader.cpp:390-612                     processor-face neighbour centres
// Geometry follows OpenFOAM's primitiveMesh face/cell centre, area and volume
// algorithm (the reference takes them from OpenFOAM, polyMeshReaderOF.cpp:125-135,
// 274-281).
#include "mesh_builder.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace lfm {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double comp(const V3 &a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

void face_geometry(const lfm_polymesh &m, std::vector<V3> &fc, std::vector<V3> &fa) {
    fc.resize(m.n_faces);
    fa.resize(m.n_faces);
    auto P = [&](int i) { return V3{m.points[3 * i], m.points[3 * i + 1], m.points[3 * i + 2]}; };
    for (int f = 0; f < m.n_faces; f++) {
        const int32_t *fp = m.face_pts + m.face_off[f];
        const int np = m.face_off[f + 1] - m.face_off[f];
        if (np == 3) {
            V3 p0 = P(fp[0]), p1 = P(fp[1]), p2 = P(fp[2]);
            fc[f] = {(1.0 / 3.0) * (p0.x + p1.x + p2.x), (1.0 / 3.0) * (p0.y + p1.y + p2.y),
                     (1.0 / 3.0) * (p0.z + p1.z + p2.z)};
            V3 n = cross(p1 - p0, p2 - p0);
            fa[f] = 0.5 * n;
            continue;
        }
        V3 est{0, 0, 0};
        for (int i = 0; i < np; i++) est = est + P(fp[i]);
        est = {est.x / np, est.y / np, est.z / np};
        V3 sumN{0, 0, 0}, sumAc{0, 0, 0};
        double sumA = 0.0;
        for (int i = 0; i < np; i++) {
            V3 a = P(fp[i]), b = P(fp[(i + 1) % np]);
            V3 c = a + b + est;
            V3 n = cross(b - a, est - a);
            double mag = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
            sumN = sumN + n;
            sumAc = {sumAc.x + mag * c.x, sumAc.y + mag * c.y, sumAc.z + mag * c.z};
            sumA += mag;
        }
        if (sumA < 1.0e-300) {
            fc[f] = est;
            fa[f] = {0, 0, 0};
        } else {
            fc[f] = {(1.0 / 3.0) * sumAc.x / sumA, (1.0 / 3.0) * sumAc.y / sumA, (1.0 / 3.0) * sumAc.z / sumA};
            fa[f] = 0.5 * sumN;
        }
    }
}

void cell_geometry(const lfm_polymesh &m, const std::vector<V3> &fc, const std::vector<V3> &fa,
                   std::vector<V3> &cc, std::vector<double> &vol) {
    const int nc = m.n_cells;
    std::vector<V3> est(nc, V3{0, 0, 0});
    std::vector<int> cnt(nc, 0);
    for (int f = 0; f < m.n_faces; f++) {
        est[m.owner[f]] = est[m.owner[f]] + fc[f];
        cnt[m.owner[f]]++;
    }
    for (int f = 0; f < m.n_internal; f++) {
        est[m.neighbour[f]] = est[m.neighbour[f]] + fc[f];
        cnt[m.neighbour[f]]++;
    }
    for (int c = 0; c < nc; c++) est[c] = {est[c].x / cnt[c], est[c].y / cnt[c], est[c].z / cnt[c]};
    cc.assign(nc, V3{0, 0, 0});
    vol.assign(nc, 0.0);
    auto add = [&](int c, int f, double pyr) {
        V3 pc{(3.0 / 4.0) * fc[f].x + (1.0 / 4.0) * est[c].x, (3.0 / 4.0) * fc[f].y + (1.0 / 4.0) * est[c].y,
              (3.0 / 4.0) * fc[f].z + (1.0 / 4.0) * est[c].z};
        cc[c] = {cc[c].x + pyr * pc.x, cc[c].y + pyr * pc.y, cc[c].z + pyr * pc.z};
        vol[c] += pyr;
    };
    for (int f = 0; f < m.n_faces; f++) {
        int o = m.owner[f];
        add(o, f, dot(fa[f], fc[f] - est[o]));
    }
    for (int f = 0; f < m.n_internal; f++) {
        int n = m.neighbour[f];
        add(n, f, dot(fa[f], est[n] - fc[f]));
    }
    for (int c = 0; c < nc; c++) {
        if (std::fabs(vol[c]) > 1.0e-300)
            cc[c] = {cc[c].x / vol[c], cc[c].y / vol[c], cc[c].z / vol[c]};
        else
            cc[c] = est[c];
        vol[c] *= (1.0 / 3.0);
    }
}

}  // namespace

Gas derive_gas(const lfm_params &prm) {
    Gas g{};
    const int D = prm.dim;
    g.R = 8.31447 / prm.mol_weight * 1000;
    g.cp = prm.cp;
    g.gamma = prm.cp / (prm.cp - g.R);
    g.gm1 = g.gamma - 1.0;
    g.R_inv = 1.0 / g.R;
    g.mu = prm.mu;
    g.pr_inv = 1.0 / prm.pr;
    g.p_inf = prm.p_inf;
    g.T_inf = prm.T_inf;
    g.rho_inf = prm.p_inf / (g.R * prm.T_inf);
    double u2 = 0;
    for (int k = 0; k < 3; k++) g.U_inf[k] = prm.U_inf[k];
    for (int k = 0; k < D; k++) u2 += prm.U_inf[k] * prm.U_inf[k];
    g.E_inf = prm.p_inf / (g.rho_inf * g.gm1) + 0.5 * u2;
    return g;
}

int rk_tables(int rk_order, double A[5], double B[5]) {
    for (int i = 0; i < 5; i++) A[i] = B[i] = 0.0;
    switch (rk_order) {
    case 0:
    case 1:
        A[0] = 0.0; B[0] = 1.0;
        break;
    case 2:
        A[0] = 0.0; A[1] = -1.0; B[0] = 1.0; B[1] = 0.5;
        break;
    case 3:
    case 4:
        A[0] = 0.0; A[1] = -756391.0 / 934407.0; A[2] = -36441873.0 / 15625000.0; A[3] = -1953125.0 / 1085297.0;
        B[0] = 8.0 / 141.0; B[1] = 6627.0 / 2000.0; B[2] = 609375.0 / 1085297.0; B[3] = 198961.0 / 526383.0;
        break;
    case 5:
        A[0] = 0.0; A[1] = -0.4178904745; A[2] = -1.192151694643; A[3] = -1.697784692471; A[4] = -1.514183444257;
        B[0] = 0.1496590219993; B[1] = 0.3792103129999; B[2] = 0.8229550293869; B[3] = 0.6994504559488;
        B[4] = 0.1530572479681;
        break;
    default:
        return -1;
    }
    // Mesh::solve runs rkOrder stages (mesh_solver.cpp:500): rkOrder 3 uses
    // the first three stages of the four-stage table, rkOrder 0 none.
    return rk_order;
}

std::string build_layout(const lfm_polymesh &m, const lfm_partition *part, const lfm_params &prm,
                         const double *p, const double *T, const double *U3, const double *alpha,
                         SolverLayout &L) {
    const int D = prm.dim;
    if (D != 2 && D != 3) return "dimension must be 2 or 3";
    if (m.n_cells <= 0 || m.n_faces <= 0) return "empty mesh";
    if (m.n_internal < 0 || m.n_internal > m.n_faces) return "n_internal out of range";
    const int nc = m.n_cells, nf = m.n_faces;
    // validate patches
    std::vector<int32_t> fpatch(nf, -1);
    for (int pch = 0; pch < m.n_patches; pch++) {
        const int s = m.patch_start[pch], n = m.patch_size[pch];
        if (s < m.n_internal || n < 0 || (long)s + n > nf) return "patch range out of bounds";
        for (int f = s; f < s + n; f++) fpatch[f] = pch;
        int role = m.patch_role[pch];
        if (role < LFM_BC_UNSET || role > LFM_BC_PROCESSOR) return "unknown patch role";
        if (role == LFM_BC_PROCESSOR) {
            if (!part || !part->face_key) return "processor patch without partition data";
            if (!m.patch_nbr || m.patch_nbr[pch] < 0 || m.patch_nbr[pch] == part->rank)
                return "processor patch without a neighbour rank";
        }
        if (role == LFM_BC_CYCLIC) {
            int nb = m.patch_nbr ? m.patch_nbr[pch] : -1;
            if (nb < 0 || nb >= m.n_patches || m.patch_size[nb] != n)
                return "cyclic patch without a partner of equal size";
        }
    }
    for (int f = m.n_internal; f < nf; f++)
        if (fpatch[f] < 0) return "boundary face outside every patch";
    if (m.n_patches < 0) return "negative patch count";
    if (m.n_points <= 0 || m.face_off[0] != 0) return "face offsets must start at 0 over a non-empty point list";
    for (int f = 0; f < nf; f++) {
        if (m.owner[f] < 0 || m.owner[f] >= nc) return "owner out of range";
        if ((long)m.face_off[f + 1] - m.face_off[f] < 3) return "face with fewer than 3 points";
    }
    for (int32_t k = 0; k < m.face_off[nf]; k++)
        if (m.face_pts[k] < 0 || m.face_pts[k] >= m.n_points) return "face point label out of range";
    for (int f = 0; f < m.n_internal; f++)
        if (m.neighbour[f] < 0 || m.neighbour[f] >= nc) return "neighbour out of range";

    std::vector<V3> fc, fa, cc;
    std::vector<double> vol;
    face_geometry(m, fc, fa);
    cell_geometry(m, fc, fa, cc, vol);

    auto is_valid = [&](int f) { return !(fpatch[f] >= 0 && m.patch_role[fpatch[f]] == LFM_BC_EMPTY); };
    auto twin = [&](int f) {
        int pch = fpatch[f];
        if (pch < 0 || m.patch_role[pch] != LFM_BC_CYCLIC) return f;
        return f - m.patch_start[pch] + m.patch_start[m.patch_nbr[pch]];
    };
    auto face_nbr = [&](int f) {
        if (f < m.n_internal) return (int)m.neighbour[f];
        int t = twin(f);
        return t == f ? -1 : (int)m.owner[t];
    };

    // cells(): owned faces ascending, then neighboured faces ascending
    std::vector<int32_t> coff(nc + 1, 0), cfaces;
    {
        std::vector<int32_t> cnt(nc, 0);
        for (int f = 0; f < nf; f++) cnt[m.owner[f]]++;
        for (int f = 0; f < m.n_internal; f++) cnt[m.neighbour[f]]++;
        for (int c = 0; c < nc; c++) coff[c + 1] = coff[c] + cnt[c];
        cfaces.resize(coff[nc]);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int f = 0; f < nf; f++) cfaces[coff[m.owner[f]] + cnt[m.owner[f]]++] = f;
        for (int f = 0; f < m.n_internal; f++) cfaces[coff[m.neighbour[f]] + cnt[m.neighbour[f]]++] = f;
    }

    // submesh split
    std::vector<int8_t> sm(nc, 1);
    if (m.cell_submesh) {
        for (int c = 0; c < nc; c++) {
            if (m.cell_submesh[c] != 0 && m.cell_submesh[c] != 1) return "cellSubmesh must be 0 or 1";
            sm[c] = (int8_t)m.cell_submesh[c];
        }
    } else {
        std::vector<uint8_t> bpt(m.n_points, 0);
        for (int pch = 0; pch < m.n_patches; pch++) {
            if (m.patch_role[pch] == LFM_BC_EMPTY) continue;
            for (int f = m.patch_start[pch]; f < m.patch_start[pch] + m.patch_size[pch]; f++)
                for (int k = m.face_off[f]; k < m.face_off[f + 1]; k++) bpt[m.face_pts[k]] = 1;
        }
        for (int c = 0; c < nc; c++) {
            bool interior = true;
            for (int i = coff[c]; i < coff[c + 1] && interior; i++) {
                int f = cfaces[i];
                for (int k = m.face_off[f]; k < m.face_off[f + 1] && interior; k++) interior = !bpt[m.face_pts[k]];
            }
            sm[c] = interior ? 1 : 0;
        }
    }

    L = SolverLayout();
    L.D = D;
    L.nc = nc;
    L.orig.resize(nc);
    std::vector<int32_t> comb(nc);
    int idx = 0;
    for (int c = 0; c < nc; c++)
        if (sm[c] == 0) { L.orig[idx] = c; comb[c] = idx++; }
    L.nsm0 = idx;
    for (int c = 0; c < nc; c++)
        if (sm[c] != 0) { L.orig[idx] = c; comb[c] = idx++; }

    // per-cell valid slots (original slot order)
    struct Slot {
        int face;       // polyMesh face id seen by this cell
        int nb;         // combined index (>= nc: ghost)
        double S[3], d[3], w;
    };
    std::vector<int32_t> soff(nc + 1, 0);
    for (int ci = 0; ci < nc; ci++) {
        int c = L.orig[ci], n = 0;
        for (int i = coff[c]; i < coff[c + 1]; i++) n += is_valid(cfaces[i]);
        if (n > kMaxFaces) return "a cell has more than 6 valid faces";
        if (n < 3) return "a cell has fewer than 3 valid faces";
        soff[ci + 1] = soff[ci] + n;
        L.max_faces = std::max(L.max_faces, n);
    }
    std::vector<Slot> slots(soff[nc]);

    // ghosts: boundary patches first, then processor patches; each patch's
    // ghosts in submesh-0 cell order, slots in order (cfd_v0.cpp:384-431)
    std::vector<int32_t> gstart(m.n_patches + 1, 0);
    {
        int acc = 0;
        for (int pass = 0; pass < 2; pass++) {
            for (int pch = 0; pch < m.n_patches; pch++) {
                int role = m.patch_role[pch];
                if ((role == LFM_BC_PROCESSOR) != (pass == 1)) continue;
                gstart[pch] = acc;
                if (role == LFM_BC_EMPTY || role == LFM_BC_CYCLIC) continue;
                for (int ci = 0; ci < nc; ci++) {
                    int c = L.orig[ci];
                    for (int i = coff[c]; i < coff[c + 1]; i++) {
                        int f = cfaces[i];
                        if (is_valid(f) && fpatch[f] == pch) {
                            if (ci >= L.nsm0) return "boundary face on an interior-submesh cell";
                            L.ghost_cell.push_back(ci);
                            L.ghost_role.push_back(role);
                            L.ghost_patch.push_back(pch);
                            acc++;
                        }
                    }
                }
            }
            if (pass == 0) L.ng_bnd = acc;
        }
        L.ng = acc;
    }
    L.rank = part ? part->rank : -1;
    L.patch_role.assign(m.patch_role, m.patch_role + m.n_patches);
    std::vector<int32_t> gcount(m.n_patches + 1, 0);
    std::vector<int32_t> ghost_slot(L.ng_bnd, -1);  // original slot of each boundary ghost's face
    for (int ci = 0; ci < nc; ci++) {
        const int c = L.orig[ci];
        int s = soff[ci];
        for (int i = coff[c]; i < coff[c + 1]; i++) {
            const int f = cfaces[i];
            if (!is_valid(f)) continue;
            Slot &sl = slots[s++];
            const int own = m.owner[f], nb = face_nbr(f);
            const int ncell = (own == c) ? nb : own;
            const double vsign = (own == c) ? 1.0 : -1.0;
            sl.face = f;
            double dP[3];
            for (int k = 0; k < D; k++) {
                sl.S[k] = vsign * comp(fa[f], k);
                dP[k] = comp(cc[c], k) - comp(fc[f], k);
            }
            if (ncell < 0) {
                double S2 = 0.0, dPS = 0.0;
                for (int k = 0; k < D; k++) S2 += sl.S[k] * sl.S[k];
                for (int k = 0; k < D; k++) dPS += dP[k] * sl.S[k];
                for (int k = 0; k < D; k++) sl.d[k] = 2.0 * std::fabs(dPS) * sl.S[k] / S2;
                const int pch = fpatch[f];
                sl.nb = nc + gstart[pch] + gcount[pch]++;
                if (sl.nb - nc < L.ng_bnd) ghost_slot[sl.nb - nc] = s - 1;
                if (m.patch_role[pch] == LFM_BC_PROCESSOR) {
                    // halo ghost; d and w are replaced by halo_import
                    const int rank = m.patch_nbr[pch];
                    size_t ni = 0;
                    while (ni < L.halo.size() && L.halo[ni].rank != rank) ni++;
                    if (ni == L.halo.size()) {
                        L.halo.emplace_back();
                        L.halo.back().rank = rank;
                    }
                    HaloNbr &h = L.halo[ni];
                    h.ghost.push_back(sl.nb - nc);
                    h.key.push_back(part->face_key[f - m.n_internal]);
                    h.cell.push_back(ci);
                    for (int k = 0; k < 3; k++) {
                        h.xf.push_back(comp(fc[f], k));
                        h.xc.push_back(comp(cc[c], k));
                    }
                }
            } else {
                double vn[3];
                for (int k = 0; k < D; k++) vn[k] = comp(cc[nb], k) - comp(cc[own], k);
                if (f >= m.n_internal) {
                    int t = twin(f);
                    for (int k = 0; k < D; k++) vn[k] += comp(fc[f], k) - comp(fc[t], k);
                }
                for (int k = 0; k < D; k++) sl.d[k] = vsign * vn[k];
                sl.nb = comb[ncell];
            }
            double wn = 0.0, dv = 0.0;
            for (int k = 0; k < D; k++) {
                double dN = dP[k] + sl.d[k];
                wn += sl.S[k] * dN;
                dv += (-sl.S[k] * dP[k] + sl.S[k] * dN);
            }
            sl.w = wn / dv;
            for (int k = D; k < 3; k++) sl.S[k] = sl.d[k] = 0.0;
        }
    }

    // (submesh, id) ordering key of reorder_faces; ghosts sort last
    auto key_sm = [&](int i) { return i >= nc ? kGhostSubmesh : (i < L.nsm0 ? 0 : 1); };
    auto key_id = [&](int i) { return i >= nc ? i - nc + L.nsm0 : (i < L.nsm0 ? i : i - L.nsm0); };

    // reorder_faces: valid slots first (original order), invalid from the back
    std::vector<int32_t> perm(slots.size());
    std::vector<uint8_t> svalid(slots.size());
    for (int ci = 0; ci < nc; ci++) {
        const int b = soff[ci], n = soff[ci + 1] - b;
        int vi = 0, ii = n - 1;
        for (int j = 0; j < n; j++) {
            const int nbi = slots[b + j].nb;
            bool ok = true;
            if (key_sm(nbi) < key_sm(ci)) ok = false;
            else if (key_sm(nbi) == key_sm(ci) && key_id(nbi) < key_id(ci)) ok = false;
            const int dst = ok ? vi++ : ii--;
            perm[b + dst] = b + j;
            svalid[b + dst] = ok;
        }
    }

    // postProcForces' view of each boundary ghost's face
    L.gf_ref_nb.assign(L.ng_bnd, 0);
    L.gf_ref_S.assign((size_t)D * L.ng_bnd, 0.0);
    L.gf_ref_d.assign((size_t)D * L.ng_bnd, 0.0);
    L.gf_own_S.assign((size_t)D * L.ng_bnd, 0.0);
    L.gf_own_d.assign((size_t)D * L.ng_bnd, 0.0);
    for (int g = 0; g < L.ng_bnd; g++) {
        if (ghost_slot[g] < 0) return "internal error: boundary ghost without a slot";
        const Slot &own = slots[ghost_slot[g]], &ref = slots[perm[ghost_slot[g]]];
        L.gf_ref_nb[g] = ref.nb;
        for (int k = 0; k < D; k++) {
            L.gf_ref_S[(size_t)k * L.ng_bnd + g] = ref.S[k];
            L.gf_ref_d[(size_t)k * L.ng_bnd + g] = ref.d[k];
            L.gf_own_S[(size_t)k * L.ng_bnd + g] = own.S[k];
            L.gf_own_d[(size_t)k * L.ng_bnd + g] = own.d[k];
        }
    }

    // faces in processing order
    std::vector<int32_t> slot_face(slots.size(), -1);
    for (int ci = 0; ci < nc; ci++)
        for (int s = soff[ci]; s < soff[ci + 1]; s++)
            if (svalid[s]) {
                const Slot &sl = slots[perm[s]];
                slot_face[perm[s]] = L.nf;
                L.fP.push_back(ci);
                L.fN.push_back(sl.nb);
                L.nf++;
            }
    const int NF = L.nf;
    if (NF >= kGhostFaceFlag) return "too many faces for one partition (2^30)";
    L.S.assign((size_t)D * NF, 0.0);
    L.d.assign((size_t)D * NF, 0.0);
    L.w.assign(NF, 0.0);
    {
        int f = 0;
        for (int ci = 0; ci < nc; ci++)
            for (int s = soff[ci]; s < soff[ci + 1]; s++)
                if (svalid[s]) {
                    const Slot &sl = slots[perm[s]];
                    for (int k = 0; k < D; k++) {
                        L.S[(size_t)k * NF + f] = sl.S[k];
                        L.d[(size_t)k * NF + f] = sl.d[k];
                    }
                    L.w[f] = sl.w;
                    f++;
                }
    }
    // the face each invalid slot belongs to: the neighbour processed it from
    // its own slot holding the same polyMesh face (or the cyclic twin)
    {
        std::vector<int32_t> by_pface(nf, -1);
        for (size_t s = 0; s < slots.size(); s++)
            if (slot_face[s] >= 0) by_pface[slots[s].face] = slot_face[s];
        for (int ci = 0; ci < nc; ci++)
            for (int s = soff[ci]; s < soff[ci + 1]; s++) {
                if (svalid[s]) continue;
                const int pf = slots[perm[s]].face;
                int fid = by_pface[pf];
                if (fid < 0) fid = by_pface[twin(pf)];
                if (fid < 0) return "internal error: unmatched face slot";
                slot_face[perm[s]] = fid;
            }
    }
    // slot lists (reordered slot order) and gather lists (processing order)
    L.s_off.assign(nc + 1, 0);
    L.s_ent.resize(slots.size());
    for (int ci = 0; ci < nc; ci++) {
        L.s_off[ci + 1] = soff[ci + 1];
        for (int s = soff[ci]; s < soff[ci + 1]; s++) {
            const int fid = slot_face[perm[s]];
            L.s_ent[s] = svalid[s] ? fid : ~fid;
        }
    }
    {
        std::vector<int32_t> cnt(nc, 0);
        for (int f = 0; f < NF; f++) {
            cnt[L.fP[f]]++;
            if (L.fN[f] < nc) cnt[L.fN[f]]++;
        }
        L.g_off.assign(nc + 1, 0);
        for (int c = 0; c < nc; c++) L.g_off[c + 1] = L.g_off[c] + cnt[c];
        L.g_ent.resize(L.g_off[nc]);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int f = 0; f < NF; f++) {
            int P = L.fP[f], N = L.fN[f];
            L.g_ent[L.g_off[P] + cnt[P]++] = (N < nc || N >= nc + L.ng_bnd) ? f : (f | kGhostFaceFlag);
            if (N < nc) L.g_ent[L.g_off[N] + cnt[N]++] = ~f;
        }
    }

    // halo: layout face of each processor ghost, ascending send lists
    if (!L.halo.empty()) {
        std::vector<int32_t> face_of_ghost(L.ng, -1);
        for (int f = 0; f < NF; f++)
            if (L.fN[f] >= nc) face_of_ghost[L.fN[f] - nc] = f;
        for (HaloNbr &h : L.halo) {
            h.face.resize(h.ghost.size());
            for (size_t j = 0; j < h.ghost.size(); j++) {
                h.face[j] = face_of_ghost[h.ghost[j]];
                if (h.face[j] < 0) return "internal error: processor ghost without a face";
            }
            h.send_cells = h.cell;
            std::sort(h.send_cells.begin(), h.send_cells.end());
            h.send_cells.erase(std::unique(h.send_cells.begin(), h.send_cells.end()), h.send_cells.end());
            std::vector<int32_t> keys = h.key;
            std::sort(keys.begin(), keys.end());
            if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
                return "duplicate face key on a processor patch";
        }
    }

    // per-cell volume, centre, sponge, initial state
    const Gas g = derive_gas(prm);
    L.vinv.resize(nc);
    L.sigma.assign(nc, 0.0);
    L.volume.resize(nc);
    L.centre.resize((size_t)3 * nc);
    L.q0.resize((size_t)(D + 2) * nc);
    for (int ci = 0; ci < nc; ci++) {
        const int c = L.orig[ci];
        L.vinv[ci] = 1.0 / vol[c];
        L.volume[ci] = vol[c];
        L.centre[3 * ci] = cc[c].x;
        L.centre[3 * ci + 1] = cc[c].y;
        L.centre[3 * ci + 2] = cc[c].z;
    }
    if (!prm.laminar) {
        // calc_VIS_Smagorinsky (cfd_v0.cpp:1594-1733) refreshes tauMC at every
        // touch with the constant of the cell whose faces are processed; the
        // last touch of a cell is in its own loop when it has a face to
        // process, else the last face that reached it from an earlier cell
        auto cst = [&](int ci) { return -2.0 * std::pow(0.16 * std::pow(1.0 / L.vinv[ci], 1. / 3.), 2); };
        L.smag.resize(nc);
        for (int ci = 0; ci < nc; ci++) {
            int last_p = -1;
            bool own = false;
            for (int e = L.g_off[ci]; e < L.g_off[ci + 1]; e++) {
                const int ent = L.g_ent[e];
                if (ent >= 0) own = true;
                else last_p = L.fP[~ent];
            }
            L.smag[ci] = own || last_p < 0 ? cst(ci) : cst(last_p);
        }
    }
    if (prm.Ls > 0) {
        if (!alpha) return "Ls > 0 needs the alpha field";
        const double sig0 = (3.0 * (1.0 - prm.mach * prm.mach) / (prm.k_sponge * prm.Ls));
        for (int ci = 0; ci < nc; ci++) {
            const double dist = alpha[L.orig[ci]];
            if (dist < prm.Ls) {
                L.sigma[ci] = sig0 * std::pow((prm.Ls - dist) / prm.Ls, 2.0);
            }
        }
    }
    for (int ci = 0; ci < nc; ci++) {
        const int c = L.orig[ci];
        double u2 = 0;
        for (int k = 0; k < D; k++) u2 += U3[3 * c + k] * U3[3 * c + k];
        const double rho = p[c] / (g.R * T[c]);
        const double E = p[c] / (rho * g.gm1) + 0.5 * u2;
        L.q0[ci] = rho;
        for (int k = 0; k < D; k++) L.q0[(size_t)(k + 1) * nc + ci] = rho * U3[3 * c + k];
        L.q0[(size_t)(D + 1) * nc + ci] = rho * E;
    }
    return "";
}

std::string halo_import(SolverLayout &L, int ni, int n, const int32_t *key, const int32_t *cell,
                        const double *disp) {
    if (ni < 0 || ni >= (int)L.halo.size()) return "unknown neighbour";
    HaloNbr &h = L.halo[ni];
    const int nh = (int)h.key.size();
    if (n != nh) return "set-up message size differs from the processor patch";
    std::vector<std::pair<int32_t, int32_t>> idx(n);
    for (int i = 0; i < n; i++) idx[i] = {key[i], i};
    std::sort(idx.begin(), idx.end());
    for (int i = 1; i < n; i++)
        if (idx[i].first == idx[i - 1].first) return "duplicate face key in the set-up message";
    // the neighbour's send list: its processor-face cells, ascending
    std::vector<int32_t> rsend(cell, cell + n);
    std::sort(rsend.begin(), rsend.end());
    rsend.erase(std::unique(rsend.begin(), rsend.end()), rsend.end());
    const int D = L.D, NF = L.nf;
    h.recv_src.assign(nh, -1);
    for (int j = 0; j < nh; j++) {
        auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(h.key[j], (int32_t)INT32_MIN));
        if (it == idx.end() || it->first != h.key[j]) return "face key missing from the set-up message";
        const int i = it->second;
        h.recv_src[j] = (int32_t)(std::lower_bound(rsend.begin(), rsend.end(), cell[i]) - rsend.begin());
        // d = (remote centre - face centre) + face centre - cell centre, owner side
        const int f = h.face[j];
        double d[3], dP[3], S[3];
        for (int k = 0; k < D; k++) {
            d[k] = disp[3 * i + k] + h.xf[3 * j + k] - h.xc[3 * j + k];
            dP[k] = h.xc[3 * j + k] - h.xf[3 * j + k];
            S[k] = L.S[(size_t)k * NF + f];
        }
        double wn = 0.0, dv = 0.0;
        for (int k = 0; k < D; k++) {
            const double dN = dP[k] + d[k];
            wn += S[k] * dN;
            dv += (-S[k] * dP[k] + S[k] * dN);
        }
        for (int k = 0; k < D; k++) L.d[(size_t)k * NF + f] = d[k];
        L.w[f] = wn / dv;
    }
    h.imported = true;
    return "";
}

}  // namespace lfm

// Patch role from OpenFOAM patch type and name (cfd_v0.cpp:983-1002).
extern "C" int lfm_patch_role(const char *type, const char *name) {
    if (!type || !name) return LFM_BC_UNSET;
    const std::string t(type), n(name);
    if (t == "empty") return LFM_BC_EMPTY;
    if (t == "cyclic") return LFM_BC_CYCLIC;
    if (t == "processor") return LFM_BC_PROCESSOR;
    if (t == "wall") return LFM_BC_WALL;
    if (t == "patch") {
        if (n == "inlet" || n == "Inlet" || n == "Inflow" || n == "inflow") return LFM_BC_INLET;
        if (n == "outlet" || n == "Outlet" || n == "Outflow" || n == "outflow") return LFM_BC_OUTLET;
    }
    return LFM_BC_UNSET;
}
