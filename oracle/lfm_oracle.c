/*
 * lfm_oracle.c -- CPU restatement of libFastMesh's ISolver time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped,
 * never called by the product).  Plain C99, double precision, one thread.
 *
 * It restates, loop for loop, the serial path of the reference:
 *   geometry        OpenFOAM primitiveMesh face/cell centres, areas, volumes
 *                   (restated from the published algorithm; OpenFOAM is not
 *                   vendored in the reference -- parity unpinned beyond the
 *                   closed-form checks in tests/test_oracle.py)
 *   submesh split   polyMeshReaderOF/polyMeshReaderOF.cpp:501-552
 *   neighbours      src/mesh_reader.cpp:163-260, 264-338, 375-456
 *   coefficients    src/cfd_v0.cpp:760-861   (init_params)
 *   ghost set-up    src/cfd_v0.cpp:346-431   (allocate_ghost_cells)
 *   face order      src/cfd_v0.cpp:278-339   (reorder_faces)
 *   RK tables       src/cfd_v0.cpp:86-128
 *   initial state   src/cfd_v0.cpp:867-967   (set_init_conditions)
 *   BC roles        src/cfd_v0.cpp:973-1005
 *   BC ghost fill   src/cfd_v0.cpp:1010-1133
 *   prepare         src/cfd_v0.cpp:1326-1377
 *   gradients       src/cfd_v0.cpp:1384-1567
 *   viscous         src/cfd_v0.cpp:1744-1860
 *   p5 polynomials  src/cfd_v0.cpp:1863-1894
 *   M2 step         src/cfd_v0.cpp:1897-2179
 *   M2AUSM step     src/cfd_v0.cpp:2186-2523
 *   M1 step         src/cfd_v0.cpp:2530-2832
 *   CFL / dt        src/cfd_v0.cpp:2887-2942
 *   helpers         api/cfdv0_solver.h:14, 252-370
 *   stage order     src/mesh_solver.cpp:474-740
 *
 * The reference runs submesh 0 (boundary cells) then submesh 1 (interior)
 * with separate loops; every loop here runs over the concatenation
 * [submesh 0 | submesh 1], which visits cells and faces in exactly the same
 * order (the per-submesh loops read no state written by a later loop of the
 * same stage except through the in-place update that the reference's face
 * ordering makes exact).
 *
 * Deviation, documented: the reference dereferences all FACE_CNT slots in
 * reorder_faces and so crashes on cells with fewer faces than the submesh
 * maximum (mixed elements); here each cell keeps its own face count, and the
 * CFL/dt loops visit the cell's own faces.
 */
#include "lfm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXF 6
#define GHOST_SM 10000 /* MAX_MPI_RANKS, api/mpi_env.h:4 */

typedef struct {
    double q[5], dq[5];
    double rho_grad[3], rhoU_grad[3][3], rhoE_grad[3], Rpsi_grad[3], c_grad[3];
    double p_grad[3], U_grad[3][3];
    double dudx[3][3], dTdx[3], tauMC[3][3], sigmaU[3];
    double vol_inv;
    double RES[5];
    int id, sm_id, is_ghost;
} solvars;

typedef struct {
    int nface;          /* valid (non-empty) faces of the cell            */
    int nvalid;         /* m_nFaceCount after reorder_faces               */
    int neigh[MAXF];    /* index into the vars array                      */
    double S[MAXF][3], d[MAXF][3], w[MAXF];
    double sigma;       /* sponge_sigma                                   */
} cellgeo;

typedef struct {
    int D, solver, nrk, minmod, adjust, have_res, laminar;
    double max_co, dt, cfl;
    int nc, ng, nsm0;
    int *orig;            /* combined index -> original cell id        */
    int *comb;            /* original cell id -> combined index        */
    solvars *v;           /* [nc + ng]                                 */
    cellgeo *g;           /* [nc]                                      */
    /* ghost bookkeeping */
    int *ghost_cell;      /* [ng] interior combined index              */
    int *ghost_role;      /* [ng]                                      */
    int *ghost_patch;     /* [ng] patch of the ghost's face              */
    int *ghost_slot;      /* [ng] slot of that face in its cell BEFORE
                             reorder_faces (boundaries[ib][i][1])        */
    int n_walls, *wall_patch;  /* m_nWallBCList, patch order           */
    double *pAVG, *pRMS;  /* [nc] postProcAverage accumulators         */
    double *centre, *volume;
    /* physics */
    double Ak[5], Bk[5];
    double gamma, gm1, R, R_inv, pr_inv, cp, mu;
    double p_inf, T_inf, U_inf[3], rho_inf, E_inf;
    double RESsum[5];
    /* processor faces (grouped by neighbour, first-appearance order) and
     * MPI ghosts v[nc + ng + ...] (one per distinct neighbour cell) */
    int nproc, nnb;
    int *pf_cell, *pf_slot, *pf_key;
    double *pf_xf, *pf_xc;
    int nb_rank[64], nb_off[65];
    int *nb_recv[64], nb_nrecv[64], nb_imported[64];
} oracle;

/* ------------------------------------------------------------------ */
/* geometry: OpenFOAM primitiveMesh restated                          */
/* ------------------------------------------------------------------ */
static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void face_geometry(const lfmo_mesh *m, double *fc, double *fa)
{
    for (int f = 0; f < m->n_faces; f++) {
        const int *fp = m->face_pts + m->face_off[f];
        int np = m->face_off[f + 1] - m->face_off[f];
        const double *P = m->points;
        if (np == 3) {
            double a[3], b[3], n[3];
            for (int k = 0; k < 3; k++) {
                fc[3 * f + k] = (1.0 / 3.0) * (P[3 * fp[0] + k] + P[3 * fp[1] + k] + P[3 * fp[2] + k]);
                a[k] = P[3 * fp[1] + k] - P[3 * fp[0] + k];
                b[k] = P[3 * fp[2] + k] - P[3 * fp[0] + k];
            }
            cross3(a, b, n);
            for (int k = 0; k < 3; k++) fa[3 * f + k] = 0.5 * n[k];
            continue;
        }
        double est[3] = {0, 0, 0};
        for (int i = 0; i < np; i++)
            for (int k = 0; k < 3; k++) est[k] += P[3 * fp[i] + k];
        for (int k = 0; k < 3; k++) est[k] /= np;
        double sumN[3] = {0, 0, 0}, sumA = 0.0, sumAc[3] = {0, 0, 0};
        for (int i = 0; i < np; i++) {
            const double *p0 = P + 3 * fp[i];
            const double *p1 = P + 3 * fp[(i + 1) % np];
            double c[3], e[3], r[3], n[3];
            for (int k = 0; k < 3; k++) {
                c[k] = p0[k] + p1[k] + est[k];
                e[k] = p1[k] - p0[k];
                r[k] = est[k] - p0[k];
            }
            cross3(e, r, n);
            double a = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; k++) {
                sumN[k] += n[k];
                sumAc[k] += a * c[k];
            }
            sumA += a;
        }
        if (sumA < 1.0e-300) {
            for (int k = 0; k < 3; k++) { fc[3 * f + k] = est[k]; fa[3 * f + k] = 0.0; }
        } else {
            for (int k = 0; k < 3; k++) {
                fc[3 * f + k] = (1.0 / 3.0) * sumAc[k] / sumA;
                fa[3 * f + k] = 0.5 * sumN[k];
            }
        }
    }
}

static void cell_geometry(const lfmo_mesh *m, const double *fc, const double *fa,
                          double *cc, double *vol)
{
    int nc = m->n_cells;
    double *est = calloc((size_t)nc * 3, sizeof(double));
    int *cnt = calloc((size_t)nc, sizeof(int));
    for (int f = 0; f < m->n_faces; f++) {
        int o = m->owner[f];
        for (int k = 0; k < 3; k++) est[3 * o + k] += fc[3 * f + k];
        cnt[o]++;
    }
    for (int f = 0; f < m->n_internal; f++) {
        int n = m->neighbour[f];
        for (int k = 0; k < 3; k++) est[3 * n + k] += fc[3 * f + k];
        cnt[n]++;
    }
    for (int c = 0; c < nc; c++)
        for (int k = 0; k < 3; k++) est[3 * c + k] /= cnt[c];
    memset(cc, 0, sizeof(double) * 3 * nc);
    memset(vol, 0, sizeof(double) * nc);
    for (int f = 0; f < m->n_faces; f++) {
        int o = m->owner[f];
        double pyr = 0.0;
        for (int k = 0; k < 3; k++) pyr += fa[3 * f + k] * (fc[3 * f + k] - est[3 * o + k]);
        for (int k = 0; k < 3; k++)
            cc[3 * o + k] += pyr * ((3.0 / 4.0) * fc[3 * f + k] + (1.0 / 4.0) * est[3 * o + k]);
        vol[o] += pyr;
    }
    for (int f = 0; f < m->n_internal; f++) {
        int n = m->neighbour[f];
        double pyr = 0.0;
        for (int k = 0; k < 3; k++) pyr += fa[3 * f + k] * (est[3 * n + k] - fc[3 * f + k]);
        for (int k = 0; k < 3; k++)
            cc[3 * n + k] += pyr * ((3.0 / 4.0) * fc[3 * f + k] + (1.0 / 4.0) * est[3 * n + k]);
        vol[n] += pyr;
    }
    for (int c = 0; c < nc; c++) {
        if (fabs(vol[c]) > 1.0e-300)
            for (int k = 0; k < 3; k++) cc[3 * c + k] /= vol[c];
        else
            for (int k = 0; k < 3; k++) cc[3 * c + k] = est[3 * c + k];
        vol[c] *= (1.0 / 3.0);
    }
    free(est);
    free(cnt);
}

/* ------------------------------------------------------------------ */
static int face_patch(const lfmo_mesh *m, int f)
{
    if (f < m->n_internal) return -1;
    for (int p = 0; p < m->n_patches; p++)
        if (f >= m->patch_start[p] && f < m->patch_start[p] + m->patch_size[p]) return p;
    return -1;
}

/* getCyclicFaceIndex, polyMeshReaderOF.cpp:226-248 */
static int cyclic_twin(const lfmo_mesh *m, int f, int p)
{
    if (p < 0 || m->patch_role[p] != LFMO_BC_CYCLIC || m->patch_nbr[p] < 0) return f;
    return f - m->patch_start[p] + m->patch_start[m->patch_nbr[p]];
}

static void seterr(char *err, int errlen, const char *msg)
{
    if (err && errlen > 0) {
        strncpy(err, msg, (size_t)errlen - 1);
        err[errlen - 1] = 0;
    }
}

void *lfmo_create(const lfmo_mesh *m, const lfmo_params *prm,
                  const double *p, const double *T, const double *U3,
                  const double *alpha, char *err, int errlen)
{
    const int D = prm->dim;
    if (D != 2 && D != 3) { seterr(err, errlen, "dimension must be 2 or 3"); return NULL; }
    for (int pch = 0; pch < m->n_patches; pch++)
        if (m->patch_role[pch] == LFMO_BC_CYCLIC) {
            int nb = m->patch_nbr[pch];
            if (nb < 0 || nb >= m->n_patches || m->patch_size[nb] != m->patch_size[pch]) {
                seterr(err, errlen, "cyclic patch without a partner of equal size");
                return NULL;
            }
        }
    if (prm->solver < 0 || prm->solver > 2) { seterr(err, errlen, "solver must be 0, 1 or 2"); return NULL; }
    if (prm->rk_order < 0 || prm->rk_order > 5) { seterr(err, errlen, "rkOrder must be 0..5"); return NULL; }

    const int nc = m->n_cells, nf = m->n_faces;
    double *fc = malloc(sizeof(double) * 3 * nf), *fa = malloc(sizeof(double) * 3 * nf);
    double *cc = malloc(sizeof(double) * 3 * nc), *vol = malloc(sizeof(double) * nc);
    face_geometry(m, fc, fa);
    cell_geometry(m, fc, fa, cc, vol);

    int *fpatch = malloc(sizeof(int) * nf);
    char *valid = malloc((size_t)nf);
    for (int f = 0; f < nf; f++) {
        fpatch[f] = face_patch(m, f);
        valid[f] = !(fpatch[f] >= 0 && m->patch_role[fpatch[f]] == LFMO_BC_EMPTY);
    }

    /* submesh split: polyMeshReaderOF.cpp:501-552 */
    int *sm = malloc(sizeof(int) * nc);
    /* cells(): owned faces ascending, then neighbour faces ascending */
    int *ncf = calloc((size_t)nc, sizeof(int));
    for (int f = 0; f < nf; f++) ncf[m->owner[f]]++;
    for (int f = 0; f < m->n_internal; f++) ncf[m->neighbour[f]]++;
    int *coff = malloc(sizeof(int) * (nc + 1));
    coff[0] = 0;
    for (int c = 0; c < nc; c++) coff[c + 1] = coff[c] + ncf[c];
    int *cfaces = malloc(sizeof(int) * coff[nc]);
    memset(ncf, 0, sizeof(int) * nc);
    for (int f = 0; f < nf; f++) { int c = m->owner[f]; cfaces[coff[c] + ncf[c]++] = f; }
    for (int f = 0; f < m->n_internal; f++) { int c = m->neighbour[f]; cfaces[coff[c] + ncf[c]++] = f; }

    if (m->cell_submesh) {
        for (int c = 0; c < nc; c++) sm[c] = m->cell_submesh[c];
    } else {
        char *bpt = calloc((size_t)m->n_points, 1);
        for (int pi = 0; pi < m->n_patches; pi++) {
            if (m->patch_role[pi] == LFMO_BC_EMPTY) continue;
            for (int f = m->patch_start[pi]; f < m->patch_start[pi] + m->patch_size[pi]; f++)
                for (int k = m->face_off[f]; k < m->face_off[f + 1]; k++) bpt[m->face_pts[k]] = 1;
        }
        for (int c = 0; c < nc; c++) {
            int interior = 1;
            for (int i = coff[c]; i < coff[c + 1] && interior; i++) {
                int f = cfaces[i];
                for (int k = m->face_off[f]; k < m->face_off[f + 1] && interior; k++)
                    interior = !bpt[m->face_pts[k]];
            }
            sm[c] = interior ? 1 : 0;
        }
        free(bpt);
    }

    oracle *o = calloc(1, sizeof(oracle));
    o->D = D;
    o->solver = prm->solver;
    o->nrk = prm->rk_order;
    o->minmod = prm->minmod;
    o->adjust = prm->adjust_dt;
    o->have_res = prm->have_res;
    o->laminar = prm->laminar;
    o->max_co = prm->max_co;
    o->dt = prm->dt;
    o->nc = nc;
    o->orig = malloc(sizeof(int) * nc);
    o->comb = malloc(sizeof(int) * nc);
    int nsm0 = 0, idx = 0;
    for (int c = 0; c < nc; c++) if (sm[c] == 0) { o->orig[idx] = c; o->comb[c] = idx++; }
    nsm0 = idx;
    for (int c = 0; c < nc; c++) if (sm[c] != 0) { o->orig[idx] = c; o->comb[c] = idx++; }
    o->nsm0 = nsm0;

    /* count ghosts: one per CELL_NONE face of a submesh-0 cell */
    int ng = 0;
    for (int c = 0; c < nc; c++)
        for (int i = coff[c]; i < coff[c + 1]; i++) {
            int f = cfaces[i];
            if (!valid[f]) continue;
            int pch = fpatch[f];
            if (pch >= 0 && m->patch_role[pch] != LFMO_BC_CYCLIC && m->patch_role[pch] != LFMO_BC_PROCESSOR) {
                if (sm[c] != 0) {
                    seterr(err, errlen, "boundary face on a submesh-1 cell");
                    goto fail;
                }
                ng++;
            }
        }
    o->ng = ng;
    /* processor faces: grouped by neighbour rank in first-appearance order
     * over (submesh-0 cell, face) -- allocate_ghost_cells, cfd_v0.cpp:437-462 */
    int nproc = 0;
    for (int ci = 0; ci < nsm0; ci++) {
        int c = o->orig[ci];
        for (int i = coff[c]; i < coff[c + 1]; i++) {
            int f = cfaces[i];
            int pch = fpatch[f];
            if (!valid[f] || pch < 0 || m->patch_role[pch] != LFMO_BC_PROCESSOR) continue;
            int r = m->patch_nbr[pch], b = 0;
            while (b < o->nnb && o->nb_rank[b] != r) b++;
            if (b == o->nnb) {
                if (o->nnb == 64) { seterr(err, errlen, "more than 64 neighbour ranks"); goto fail; }
                o->nb_rank[o->nnb++] = r;
                o->nb_off[b + 1] = 0;
            }
            o->nb_off[b + 1]++;
            nproc++;
        }
    }
    for (int c = 0; c < nc; c++)
        if (sm[c] != 0)
            for (int i = coff[c]; i < coff[c + 1]; i++) {
                int pch = fpatch[cfaces[i]];
                if (pch >= 0 && m->patch_role[pch] == LFMO_BC_PROCESSOR) {
                    seterr(err, errlen, "processor face on a submesh-1 cell");
                    goto fail;
                }
            }
    if (nproc && (!m->face_key || m->rank < 0)) { seterr(err, errlen, "processor patch without partition data"); goto fail; }
    o->nb_off[0] = 0;
    for (int b = 0; b < o->nnb; b++) o->nb_off[b + 1] += o->nb_off[b];
    o->nproc = nproc;
    o->pf_cell = malloc(sizeof(int) * (nproc + 1));
    o->pf_slot = malloc(sizeof(int) * (nproc + 1));
    o->pf_key = malloc(sizeof(int) * (nproc + 1));
    o->pf_xf = malloc(sizeof(double) * 3 * (nproc + 1));
    o->pf_xc = malloc(sizeof(double) * 3 * (nproc + 1));
    o->v = calloc((size_t)(nc + ng + nproc), sizeof(solvars));
    o->g = calloc((size_t)nc, sizeof(cellgeo));
    o->ghost_cell = malloc(sizeof(int) * (ng + 1));
    o->ghost_role = malloc(sizeof(int) * (ng + 1));
    o->ghost_patch = malloc(sizeof(int) * (ng + 1));
    o->ghost_slot = malloc(sizeof(int) * (ng + 1));
    o->wall_patch = malloc(sizeof(int) * (m->n_patches + 1));
    o->n_walls = 0;
    for (int pch = 0; pch < m->n_patches; pch++)
        if (m->patch_role[pch] == LFMO_BC_WALL) o->wall_patch[o->n_walls++] = pch;
    o->centre = cc;
    o->volume = vol;

    for (int ci = 0; ci < nc; ci++) {
        o->v[ci].sm_id = ci < nsm0 ? 0 : 1;
        o->v[ci].id = ci < nsm0 ? ci : ci - nsm0;
        o->v[ci].vol_inv = 1.0 / vol[o->orig[ci]];
    }

    /* ghosts grouped by patch (boundaries[ibound]), cells in submesh-0 order,
     * slots in original (pre-reorder) order: cfd_v0.cpp:384-431 */
    {
        int gi = 0;
        for (int pch = 0; pch < m->n_patches; pch++) {
            if (m->patch_role[pch] == LFMO_BC_EMPTY || m->patch_role[pch] == LFMO_BC_CYCLIC ||
                m->patch_role[pch] == LFMO_BC_PROCESSOR) continue;
            for (int ci = 0; ci < nsm0; ci++) {
                int c = o->orig[ci];
                for (int i = coff[c]; i < coff[c + 1]; i++) {
                    int f = cfaces[i];
                    if (!valid[f] || fpatch[f] != pch) continue;
                    o->ghost_cell[gi] = ci;
                    o->ghost_role[gi] = m->patch_role[pch];
                    o->v[nc + gi].sm_id = GHOST_SM;
                    o->v[nc + gi].id = nsm0 + gi;
                    gi++;
                }
            }
        }
    }

    /* per-cell slots: readMeshCells + setParams + calculateVNeighbour + init_params */
    {
        /* per-patch running counter to find the ghost of each boundary face */
        int *gstart = calloc((size_t)m->n_patches + 1, sizeof(int));
        int *gcount = calloc((size_t)m->n_patches + 1, sizeof(int));
        int acc = 0;
        for (int pch = 0; pch < m->n_patches; pch++) {
            gstart[pch] = acc;
            if (m->patch_role[pch] == LFMO_BC_EMPTY || m->patch_role[pch] == LFMO_BC_CYCLIC ||
                m->patch_role[pch] == LFMO_BC_PROCESSOR) continue;
            for (int ci = 0; ci < nsm0; ci++) {
                int c = o->orig[ci];
                for (int i = coff[c]; i < coff[c + 1]; i++)
                    if (valid[cfaces[i]] && fpatch[cfaces[i]] == pch) acc++;
            }
        }
        int pcount[64] = {0};
        for (int ci = 0; ci < nc; ci++) {
            int c = o->orig[ci];
            cellgeo *G = &o->g[ci];
            int s = 0;
            double dP[MAXF][3];
            int ghostslot[MAXF];
            for (int i = coff[c]; i < coff[c + 1]; i++) {
                int f = cfaces[i];
                if (!valid[f]) continue;
                if (s >= MAXF) { seterr(err, errlen, "more than 6 valid faces in a cell"); free(gstart); free(gcount); goto fail; }
                int own = m->owner[f];
                int pch = fpatch[f];
                int nb;
                if (f < m->n_internal) nb = m->neighbour[f];
                else {
                    int tw = cyclic_twin(m, f, pch);
                    nb = (tw == f) ? -1 : m->owner[tw];
                }
                int ncell = (own == c) ? nb : own;
                double vsign = (own == c) ? 1.0 : -1.0;
                for (int k = 0; k < D; k++) G->S[s][k] = vsign * fa[3 * f + k];
                for (int k = 0; k < D; k++) dP[s][k] = cc[3 * c + k] - fc[3 * f + k];
                ghostslot[s] = 0;
                if (pch >= 0 && m->patch_role[pch] == LFMO_BC_PROCESSOR) {
                    /* CELL_MPI: neighbour is an MPI ghost (valid in reorder_faces,
                     * sm_id MAX_MPI_RANKS); d and w are set when the neighbour's
                     * centres arrive (lfmo_halo_import) */
                    int b = 0;
                    while (o->nb_rank[b] != m->patch_nbr[pch]) b++;
                    int k = o->nb_off[b] + pcount[b]++;
                    o->pf_cell[k] = ci;
                    o->pf_key[k] = m->face_key[f - m->n_internal];
                    for (int q = 0; q < 3; q++) {
                        o->pf_xf[3 * k + q] = fc[3 * f + q];
                        o->pf_xc[3 * k + q] = cc[3 * c + q];
                    }
                    for (int q = 0; q < D; q++) G->d[s][q] = 0.0;
                    G->neigh[s] = nc + ng + k;
                    o->v[nc + ng + k].sm_id = GHOST_SM;
                    o->v[nc + ng + k].id = k;
                } else if (ncell < 0) {
                    /* CELL_NONE: mirror, cfd_v0.cpp:812-824 */
                    double S2 = 0.0, dPS = 0.0;
                    for (int k = 0; k < D; k++) S2 += G->S[s][k] * G->S[s][k];
                    for (int k = 0; k < D; k++) dPS += dP[s][k] * G->S[s][k];
                    for (int k = 0; k < D; k++) G->d[s][k] = 2.0 * fabs(dPS) * G->S[s][k] / S2;
                    G->neigh[s] = nc + gstart[pch] + gcount[pch];
                    o->ghost_patch[gstart[pch] + gcount[pch]] = pch;
                    o->ghost_slot[gstart[pch] + gcount[pch]] = s;
                    gcount[pch]++;
                    ghostslot[s] = 1;
                } else {
                    /* v_neigh: mesh_reader.cpp:421-442 (owner-relative), times v_sign */
                    double vn[3];
                    int ow = own, ne = nb;
                    for (int k = 0; k < D; k++) vn[k] = cc[3 * ne + k] - cc[3 * ow + k];
                    if (f >= m->n_internal) {
                        int tw = cyclic_twin(m, f, pch);
                        for (int k = 0; k < D; k++) vn[k] += fc[3 * f + k] - fc[3 * tw + k];
                    }
                    for (int k = 0; k < D; k++) G->d[s][k] = vsign * vn[k];
                    G->neigh[s] = o->comb[ncell];
                }
                s++;
            }
            (void)ghostslot;
            G->nface = s;
            for (int j = 0; j < s; j++) {
                double dN[3], w = 0.0, div = 0.0;
                for (int k = 0; k < D; k++) dN[k] = dP[j][k] + G->d[j][k];
                for (int k = 0; k < D; k++) {
                    w += G->S[j][k] * dN[k];
                    div += (-G->S[j][k] * dP[j][k] + G->S[j][k] * dN[k]);
                }
                G->w[j] = w / div;
            }
        }
        free(gstart);
        free(gcount);
    }

    /* reorder_faces, cfd_v0.cpp:278-339 */
    for (int ci = 0; ci < nc; ci++) {
        cellgeo *G = &o->g[ci];
        cellgeo T2 = *G;
        int vi = 0, ii = G->nface - 1;
        solvars *cell = &o->v[ci];
        for (int j = 0; j < G->nface; j++) {
            solvars *nb = &o->v[G->neigh[j]];
            int ok = 1;
            if (nb->sm_id < cell->sm_id) ok = 0;
            else if (nb->sm_id == cell->sm_id && nb->id < cell->id) ok = 0;
            int dst = ok ? vi++ : ii--;
            T2.neigh[dst] = G->neigh[j];
            T2.w[dst] = G->w[j];
            for (int k = 0; k < 3; k++) { T2.S[dst][k] = G->S[j][k]; T2.d[dst][k] = G->d[j][k]; }
        }
        T2.nvalid = vi;
        *G = T2;
    }

    /* slot of each processor face after the reorder */
    for (int ci = 0; ci < nc; ci++)
        for (int j = 0; j < o->g[ci].nface; j++) {
            int n = o->g[ci].neigh[j];
            if (n >= nc + ng) o->pf_slot[n - nc - ng] = j;
        }

    /* RK tables, cfd_v0.cpp:86-128 */
    switch (prm->rk_order) {
    case 0: case 1:
        o->Ak[0] = 0.0; o->Bk[0] = 1.0; break;
    case 2:
        o->Ak[0] = 0.0; o->Ak[1] = -1.0; o->Bk[0] = 1.0; o->Bk[1] = 0.5; break;
    case 3: case 4:
        o->Ak[0] = 0.0; o->Ak[1] = -756391.0 / 934407.0; o->Ak[2] = -36441873.0 / 15625000.0;
        o->Ak[3] = -1953125.0 / 1085297.0;
        o->Bk[0] = 8.0 / 141.0; o->Bk[1] = 6627.0 / 2000.0; o->Bk[2] = 609375.0 / 1085297.0;
        o->Bk[3] = 198961.0 / 526383.0; break;
    case 5:
        o->Ak[0] = 0.0; o->Ak[1] = -0.4178904745; o->Ak[2] = -1.192151694643;
        o->Ak[3] = -1.697784692471; o->Ak[4] = -1.514183444257;
        o->Bk[0] = 0.1496590219993; o->Bk[1] = 0.3792103129999; o->Bk[2] = 0.8229550293869;
        o->Bk[3] = 0.6994504559488; o->Bk[4] = 0.1530572479681; break;
    }

    /* set_init_conditions, cfd_v0.cpp:880-965 */
    o->p_inf = prm->p_inf;
    o->T_inf = prm->T_inf;
    for (int k = 0; k < 3; k++) o->U_inf[k] = prm->U_inf[k];
    o->mu = prm->mu;
    o->cp = prm->cp;
    o->R = 8.31447 / prm->mol_weight * 1000;
    {
        double Cv = o->cp - o->R;
        o->gamma = o->cp / Cv;
    }
    o->rho_inf = o->p_inf / (o->R * o->T_inf);
    o->gm1 = o->gamma - 1.0;
    {
        double u2 = 0;
        for (int k = 0; k < D; k++) u2 += o->U_inf[k] * o->U_inf[k];
        o->E_inf = o->p_inf / (o->rho_inf * o->gm1) + 0.5 * u2;
    }
    o->R_inv = 1.0 / o->R;
    o->pr_inv = 1.0 / prm->pr;

    if (prm->Ls > 0) {
        double sig0 = (3.0 * (1.0 - prm->mach * prm->mach) / (prm->k_sponge * prm->Ls));
        for (int ci = 0; ci < nc; ci++) {
            double dist = alpha[o->orig[ci]];
            if (dist < prm->Ls) o->g[ci].sigma = sig0 * pow((prm->Ls - dist) / prm->Ls, 2.0);
        }
    }
    for (int ci = 0; ci < nc; ci++) {
        int c = o->orig[ci];
        double u2 = 0;
        for (int k = 0; k < D; k++) u2 += U3[3 * c + k] * U3[3 * c + k];
        double rho = p[c] / (o->R * T[c]);
        double E = p[c] / (rho * o->gm1) + 0.5 * u2;
        o->v[ci].q[0] = rho;
        for (int k = 0; k < D; k++) o->v[ci].q[k + 1] = rho * U3[3 * c + k];
        o->v[ci].q[D + 1] = rho * E;
    }

    free(fc); free(fa); free(fpatch); free(valid); free(sm); free(ncf); free(coff); free(cfaces);
    return o;

fail:
    free(fc); free(fa); free(fpatch); free(valid); free(sm); free(ncf); free(coff); free(cfaces);
    if (o) {
        free(o->orig); free(o->comb); free(o->v); free(o->g); free(o->ghost_cell); free(o->ghost_role);
        free(o->ghost_patch); free(o->ghost_slot); free(o->wall_patch);
        free(o);
    }
    free(cc); free(vol);
    return NULL;
}

void lfmo_destroy(void *h)
{
    oracle *o = h;
    if (!o) return;
    free(o->pf_cell); free(o->pf_slot); free(o->pf_key); free(o->pf_xf); free(o->pf_xc);
    for (int b = 0; b < o->nnb; b++) free(o->nb_recv[b]);
    free(o->orig); free(o->comb); free(o->v); free(o->g); free(o->ghost_cell); free(o->ghost_role);
    free(o->ghost_patch); free(o->ghost_slot); free(o->wall_patch); free(o->pAVG); free(o->pRMS);
    free(o->centre); free(o->volume);
    free(o);
}

/* ------------------------------------------------------------------ */
/* helpers, api/cfdv0_solver.h                                        */
/* ------------------------------------------------------------------ */
#define INTERP_LINEAR(w, a, b) ((w) * (a) + (1.0 - (w)) * (b))

static double compute_Rpsi(const oracle *o, const double *q)
{
    const int D = o->D;
    double rho_inv = 1.0 / q[0];
    double s = q[1] * q[1];
    for (int k = 1; k < D; k++) s += q[k + 1] * q[k + 1];
    return o->gm1 * (q[D + 1] - 0.5 * s * rho_inv) * rho_inv;
}

static double dotD(int D, const double *a, const double *b)
{
    double s = a[0] * b[0];
    for (int k = 1; k < D; k++) s += a[k] * b[k];
    return s;
}

static double magD(int D, const double *a) { return sqrt(dotD(D, a, a)); }

static double sgn(double a) { return a < 0.0 ? -1.0 : 1.0; }

static double calc_r(int D, double phiP, double phiN, const double *grad, const double *d)
{
    double gradf = phiN - phiP + 1.0e-30;
    double gradcf = 0.0;
    for (int k = 0; k < D; k++) gradcf += d[k] * grad[k];
    if (fabs(gradcf) >= 1000.0 * fabs(gradf)) return 2.0 * 1000.0 * sgn(gradcf) * sgn(gradf) - 1.0;
    return 2.0 * (gradcf / gradf) - 1.0;
}

static double interp_minmod(int D, double cphi, double aphi, const double *grad, const double *d,
                            double wl, double flux)
{
    double r = calc_r(D, cphi, aphi, grad, d);
    double m1 = (1.0 < r) ? 1.0 : r;       /* std::min(r, ONE) */
    double lim = (m1 < 0.0) ? 0.0 : m1;    /* std::max(., ZERO) */
    double w = lim * wl + (1.0 - lim) * (1.0 + flux) * 0.5;
    return w * cphi + (1.0 - w) * aphi;
}

static double p5Pos(double M, double alpha)
{
    double M2Pos = 0.25 * (M + 1.0) * (M + 1.0);
    double M2Neg = -0.25 * (M - 1.0) * (M - 1.0);
    double M1Pos = 0.5 * (M + fabs(M));
    if (fabs(M) < 1) return M2Pos * ((2.0 - M) - 16.0 * alpha * M * M2Neg);
    return M1Pos / M;
}

static double p5Neg(double M, double alpha)
{
    double M2Pos = 0.25 * (M + 1.0) * (M + 1.0);
    double M2Neg = -0.25 * (M - 1.0) * (M - 1.0);
    double M1Neg = 0.5 * (M - fabs(M));
    if (fabs(M) < 1) return M2Neg * ((-2.0 - M) + 16.0 * alpha * M * M2Pos);
    return M1Neg / M;
}

/* ------------------------------------------------------------------ */
/* stage pieces                                                        */
/* ------------------------------------------------------------------ */
static void prepare_for_timestep(oracle *o)
{
    for (int c = 0; c < o->nc; c++)
        for (int k = 0; k < o->D + 2; k++) o->v[c].dq[k] = o->v[c].RES[k] = 0.0;
}

static void prepare_for_RKstep(oracle *o, int rk)
{
    for (int c = 0; c < o->nc; c++) {
        solvars *v = &o->v[c];
        for (int k = 0; k < o->D + 2; k++) v->dq[k] *= o->Ak[rk];
        memset(v->rho_grad, 0, sizeof v->rho_grad);
        memset(v->rhoU_grad, 0, sizeof v->rhoU_grad);
        memset(v->rhoE_grad, 0, sizeof v->rhoE_grad);
        memset(v->Rpsi_grad, 0, sizeof v->Rpsi_grad);
        memset(v->c_grad, 0, sizeof v->c_grad);
        memset(v->dudx, 0, sizeof v->dudx);
        memset(v->dTdx, 0, sizeof v->dTdx);
        memset(v->tauMC, 0, sizeof v->tauMC);
        if (o->solver == 2) {
            memset(v->p_grad, 0, sizeof v->p_grad);
            memset(v->U_grad, 0, sizeof v->U_grad);
        }
    }
    for (int g = 0; g < o->ng; g++) o->v[o->nc + g].is_ghost = 1;
}

static void set_boundary_conditions(oracle *o)
{
    const int D = o->D;
    /* the reference walks the wall, inlet, outlet lists in that order; the
     * fills of distinct ghosts are independent, so one pass is equivalent */
    for (int g = 0; g < o->ng; g++) {
        solvars *gh = &o->v[o->nc + g];
        const solvars *in = &o->v[o->ghost_cell[g]];
        int role = o->ghost_role[g];
        if (role == LFMO_BC_WALL) {
            gh->q[0] = in->q[0];
            for (int k = 0; k < D; k++) gh->q[k + 1] = -in->q[k + 1];
            gh->q[D + 1] = in->q[D + 1];
            for (int i = 0; i < D; i++) {
                gh->rho_grad[i] = -in->rho_grad[i];
                for (int j = 0; j < D; j++) {
                    gh->rhoU_grad[i][j] = in->rhoU_grad[i][j];
                    gh->dudx[i][j] = in->dudx[i][j];
                    gh->tauMC[i][j] = in->tauMC[i][j];
                }
                gh->rhoE_grad[i] = in->rhoE_grad[i];
                gh->Rpsi_grad[i] = -in->Rpsi_grad[i];
                gh->c_grad[i] = -in->c_grad[i];
                gh->dTdx[i] = -in->dTdx[i];
            }
        } else if (role == LFMO_BC_INLET || role == LFMO_BC_OUTLET) {
            double rhoInt = in->q[0], Uint[3], u2i = 0.0;
            for (int k = 0; k < D; k++) { Uint[k] = in->q[k + 1] / rhoInt; u2i += Uint[k] * Uint[k]; }
            double pInt = (in->q[D + 1] - 0.5 * u2i * rhoInt) * o->gm1;
            double rhoExt, pExt, Uext[3], u2e = 0.0;
            if (role == LFMO_BC_INLET) {
                double TInt = pInt * o->R_inv / rhoInt;
                double TExt = o->T_inf;
                rhoExt = rhoInt * TInt / TExt;
                for (int k = 0; k < D; k++) { Uext[k] = o->U_inf[k]; u2e += Uext[k] * Uext[k]; }
                pExt = pInt;
            } else {
                pExt = o->p_inf;
                rhoExt = rhoInt * pExt / pInt;
                for (int k = 0; k < D; k++) { Uext[k] = Uint[k]; u2e += Uext[k] * Uext[k]; }
            }
            double EExt = pExt / (rhoExt * o->gm1) + 0.5 * u2e;
            gh->q[0] = rhoExt;
            for (int k = 0; k < D; k++) gh->q[k + 1] = rhoExt * Uext[k];
            gh->q[D + 1] = rhoExt * EExt;
            for (int i = 0; i < D; i++) {
                gh->rho_grad[i] = -in->rho_grad[i];
                for (int j = 0; j < D; j++) {
                    gh->rhoU_grad[i][j] = -in->rhoU_grad[i][j];
                    gh->dudx[i][j] = -in->dudx[i][j];
                    gh->tauMC[i][j] = -in->tauMC[i][j];
                }
                gh->rhoE_grad[i] = -in->rhoE_grad[i];
                gh->Rpsi_grad[i] = -in->Rpsi_grad[i];
                gh->c_grad[i] = -in->c_grad[i];
                gh->dTdx[i] = -in->dTdx[i];
            }
        }
        /* LFMO_BC_UNSET: untouched, ghost q stays 0 */
    }
}

static void calc_gradients(oracle *o)
{
    const int D = o->D;
    for (int c = 0; c < o->nc; c++) {
        cellgeo *G = &o->g[c];
        solvars *vars = &o->v[c];
        double cR = compute_Rpsi(o, vars->q);
        double cc = sqrt(o->gamma * cR);
        for (int f = 0; f < G->nvalid; f++) {
            solvars *nb = &o->v[G->neigh[f]];
            double aR = compute_Rpsi(o, nb->q);
            double ac = sqrt(o->gamma * aR);
            double w = G->w[f];
            double rho = INTERP_LINEAR(w, vars->q[0], nb->q[0]);
            double rhoU[3];
            for (int k = 0; k < D; k++) rhoU[k] = INTERP_LINEAR(w, vars->q[k + 1], nb->q[k + 1]);
            double rhoE = INTERP_LINEAR(w, vars->q[D + 1], nb->q[D + 1]);
            double Rp = INTERP_LINEAR(w, cR, aR);
            double cf = INTERP_LINEAR(w, cc, ac);
            double cs[3], as[3];
            for (int k = 0; k < D; k++) { cs[k] = G->S[f][k] * vars->vol_inv; as[k] = -G->S[f][k] * nb->vol_inv; }
            for (int i = 0; i < D; i++) {
                vars->rho_grad[i] += rho * cs[i];
                for (int j = 0; j < D; j++) vars->rhoU_grad[j][i] += rhoU[j] * cs[i];
                vars->rhoE_grad[i] += rhoE * cs[i];
                vars->Rpsi_grad[i] += Rp * cs[i];
                vars->c_grad[i] += cf * cs[i];
                nb->rho_grad[i] += rho * as[i];
                for (int j = 0; j < D; j++) nb->rhoU_grad[j][i] += rhoU[j] * as[i];
                nb->rhoE_grad[i] += rhoE * as[i];
                nb->Rpsi_grad[i] += Rp * as[i];
                nb->c_grad[i] += cf * as[i];
            }
        }
    }
}

static void calc_gradients_M2AUSM(oracle *o)
{
    const int D = o->D;
    for (int c = 0; c < o->nc; c++) {
        cellgeo *G = &o->g[c];
        solvars *vars = &o->v[c];
        double cR = compute_Rpsi(o, vars->q);
        double cc = sqrt(o->gamma * cR);
        for (int f = 0; f < G->nvalid; f++) {
            solvars *nb = &o->v[G->neigh[f]];
            double aR = compute_Rpsi(o, nb->q);
            double ac = sqrt(o->gamma * aR);
            double w = G->w[f];
            double UU[3], rhoU[3];
            for (int k = 0; k < D; k++)
                UU[k] = INTERP_LINEAR(w, vars->q[k + 1] / vars->q[0], nb->q[k + 1] / nb->q[0]);
            double rho = INTERP_LINEAR(w, vars->q[0], nb->q[0]);
            for (int k = 0; k < D; k++) rhoU[k] = INTERP_LINEAR(w, vars->q[k + 1], nb->q[k + 1]);
            double rhoE = INTERP_LINEAR(w, vars->q[D + 1], nb->q[D + 1]);
            double Rp = INTERP_LINEAR(w, cR, aR);
            double cf = INTERP_LINEAR(w, cc, ac);
            double cs[3], as[3];
            for (int k = 0; k < D; k++) { cs[k] = G->S[f][k] * vars->vol_inv; as[k] = -G->S[f][k] * nb->vol_inv; }
            double r = vars->q[0], E = vars->q[D + 1] / r, vm = 0.0;
            for (int k = 0; k < D; k++) vm += (vars->q[k + 1] / r) * (vars->q[k + 1] / r);
            double cp_ = r * o->gm1 * (E - 0.5 * vm);
            r = nb->q[0]; E = nb->q[D + 1] / r; vm = 0.0;
            for (int k = 0; k < D; k++) vm += (nb->q[k + 1] / r) * (nb->q[k + 1] / r);
            double np_ = r * o->gm1 * (E - 0.5 * vm);
            double pf = INTERP_LINEAR(w, cp_, np_);
            for (int i = 0; i < D; i++) {
                vars->p_grad[i] += pf * cs[i];
                nb->p_grad[i] += pf * as[i];
                for (int j = 0; j < D; j++) {
                    vars->U_grad[j][i] += UU[j] * cs[i];
                    nb->U_grad[j][i] += UU[j] * as[i];
                }
                vars->rho_grad[i] += rho * cs[i];
                nb->rho_grad[i] += rho * as[i];
                for (int j = 0; j < D; j++) {
                    vars->rhoU_grad[j][i] += rhoU[j] * cs[i];
                    nb->rhoU_grad[j][i] += rhoU[j] * as[i];
                }
                vars->rhoE_grad[i] += rhoE * cs[i];
                nb->rhoE_grad[i] += rhoE * as[i];
                vars->Rpsi_grad[i] += Rp * cs[i];
                nb->Rpsi_grad[i] += Rp * as[i];
                vars->c_grad[i] += cf * cs[i];
                nb->c_grad[i] += cf * as[i];
            }
        }
    }
}

/* tau / sigmaU / tauMC refresh of one cell after a touch, cfd_v0.cpp:1813-1834 */
static void vis_refresh(const oracle *o, solvars *v, const double *U)
{
    const int D = o->D;
    const double mu = o->mu;
    double tau[3][3];
    for (int a = 0; a < D; a++) {
        tau[a][a] = 2.0 * v->dudx[a][a];
        for (int b1 = a + 1; b1 < D + a; b1++) {
            int b = b1 % D;
            tau[a][b] = mu * (v->dudx[a][b] + v->dudx[b][a]);
            tau[a][a] -= v->dudx[b][b];
        }
        tau[a][a] *= 2.0 / 3.0 * mu;
    }
    double diag = 0;
    for (int a = 0; a < D; a++) diag -= v->dudx[a][a];
    diag *= mu * 2.0 / 3.0;
    for (int i = 0; i < D; i++) {
        v->sigmaU[i] = dotD(D, U, tau[i]);
        for (int j = 0; j < D; j++) v->tauMC[i][j] = mu * v->dudx[j][i];
        v->tauMC[i][i] += diag;
    }
}

static void calc_VIS(oracle *o)
{
    const int D = o->D;
    for (int c = 0; c < o->nc; c++) {
        cellgeo *G = &o->g[c];
        solvars *vars = &o->v[c];
        double cR = compute_Rpsi(o, vars->q);
        for (int f = 0; f < G->nvalid; f++) {
            solvars *nb = &o->v[G->neigh[f]];
            double aR = compute_Rpsi(o, nb->q);
            double cri = 1.0 / vars->q[0], ari = 1.0 / nb->q[0];
            double w = G->w[f];
            double fU[3], cs[3], as[3], cU[3], aU[3];
            for (int k = 0; k < D; k++) fU[k] = INTERP_LINEAR(w, vars->q[k + 1] * cri, nb->q[k + 1] * ari);
            double fR = INTERP_LINEAR(w, cR, aR);
            double fT = fR * o->R_inv;
            for (int k = 0; k < D; k++) { cs[k] = G->S[f][k] * vars->vol_inv; as[k] = -G->S[f][k] * nb->vol_inv; }
            for (int i = 0; i < D; i++)
                for (int j = 0; j < D; j++) {
                    vars->dudx[i][j] += fU[i] * cs[j];
                    nb->dudx[i][j] += fU[i] * as[j];
                }
            for (int i = 0; i < D; i++) {
                vars->dTdx[i] += fT * cs[i];
                nb->dTdx[i] += fT * as[i];
            }
            for (int k = 0; k < D; k++) { cU[k] = vars->q[k + 1] / vars->q[0]; aU[k] = nb->q[k + 1] / nb->q[0]; }
            vis_refresh(o, vars, cU);
            vis_refresh(o, nb, aU);
        }
    }
}

/* calc_VIS_Smagorinsky (cfd_v0.cpp:1574-1737): calc_VIS's accumulation, and
 * at every touch of a cell the laminar refresh plus the constant-coefficient
 * Smagorinsky stress with the constant -2 (0.16 V^(1/3))^2 of the cell whose
 * faces are being processed (also when refreshing the neighbour). */
static void vis_refresh_smag(const oracle *o, solvars *v, const double *U, double smag)
{
    const int D = o->D;
    vis_refresh(o, v, U);
    double strain = 0.0, divu = 0.0, sij;
    for (int i = 0; i < D; i++) {
        for (int j = 0; j < D; j++) {
            sij = 0.5 * (v->dudx[i][j] + v->dudx[j][i]);
            strain += pow(sij, 2);
        }
        divu += v->dudx[i][i];
    }
    strain = sqrt(2.0 * strain);
    for (int i = 0; i < D; i++) {
        for (int j = 0; j < D; j++) {
            sij = 0.5 * (v->dudx[i][j] + v->dudx[j][i]);
            v->tauMC[i][j] -= smag * strain * sij;
        }
        v->tauMC[i][i] += smag * strain * divu / 3.0;
    }
}

static void calc_VIS_Smagorinsky(oracle *o)
{
    const int D = o->D;
    for (int c = 0; c < o->nc; c++) {
        cellgeo *G = &o->g[c];
        solvars *vars = &o->v[c];
        double cR = compute_Rpsi(o, vars->q);
        double smag = -2.0 * pow(0.16 * pow(1.0 / vars->vol_inv, 1. / 3.), 2);
        for (int f = 0; f < G->nvalid; f++) {
            solvars *nb = &o->v[G->neigh[f]];
            double aR = compute_Rpsi(o, nb->q);
            double cri = 1.0 / vars->q[0], ari = 1.0 / nb->q[0];
            double w = G->w[f];
            double fU[3], cs[3], as[3], cU[3], aU[3];
            for (int k = 0; k < D; k++) fU[k] = INTERP_LINEAR(w, vars->q[k + 1] * cri, nb->q[k + 1] * ari);
            double fR = INTERP_LINEAR(w, cR, aR);
            double fT = fR * o->R_inv;
            for (int k = 0; k < D; k++) { cs[k] = G->S[f][k] * vars->vol_inv; as[k] = -G->S[f][k] * nb->vol_inv; }
            for (int i = 0; i < D; i++)
                for (int j = 0; j < D; j++) {
                    vars->dudx[i][j] += fU[i] * cs[j];
                    nb->dudx[i][j] += fU[i] * as[j];
                }
            for (int i = 0; i < D; i++) {
                vars->dTdx[i] += fT * cs[i];
                nb->dTdx[i] += fT * as[i];
            }
            for (int k = 0; k < D; k++) { cU[k] = vars->q[k + 1] / vars->q[0]; aU[k] = nb->q[k + 1] / nb->q[0]; }
            vis_refresh_smag(o, vars, cU, smag);
            vis_refresh_smag(o, nb, aU, smag);
        }
    }
}

/* viscous face contribution shared by M1/M2/M2AUSM (cfd_v0.cpp:2682-2790) */
static void viscous_face(const oracle *o, const cellgeo *G, int f, const solvars *cell, const solvars *nb,
                         const double *cU, const double *aU, double cT, double aT, double S_mag, double *rhs)
{
    const int D = o->D;
    const double mu = o->mu;
    const double w = G->w[f];
    const double *S = G->S[f], *d = G->d[f];
    double d_mag = magD(D, d);
    double dmi = 1.0 / d_mag;
    double dn[3], dudx[3][3], dTdx[3], tau[3][3], tauMC[3][3], divT[3], delta[3], K[3], lapU[3], sU[3] = {0, 0, 0};
    for (int k = 0; k < D; k++) dn[k] = d[k] * dmi;
    for (int i = 0; i < D; i++) {
        for (int j = 0; j < D; j++) dudx[i][j] = INTERP_LINEAR(w, cell->dudx[i][j], nb->dudx[i][j]);
        dTdx[i] = INTERP_LINEAR(w, cell->dTdx[i], nb->dTdx[i]);
    }
    if (nb->is_ghost)
        for (int i = 0; i < D; i++) {
            for (int j = 0; j < D; j++) dudx[i][j] = (aU[i] - cU[i]) * dn[j] * dmi;
            dTdx[i] = (aT - cT) * dn[i] * dmi;
        }
    for (int a = 0; a < D; a++) {
        tau[a][a] = 2.0 * dudx[a][a];
        for (int b1 = a + 1; b1 < D + a; b1++) {
            int b = b1 % D;
            tau[a][b] = mu * (dudx[a][b] + dudx[b][a]);
            tau[a][a] -= dudx[b][b];
        }
        tau[a][a] *= 2.0 / 3.0 * mu;
    }
    double k_ = o->cp * mu * o->pr_inv;
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) tauMC[i][j] = INTERP_LINEAR(w, cell->tauMC[i][j], nb->tauMC[i][j]);
    if (nb->is_ghost) {
        double diag = 0;
        for (int a = 0; a < D; a++) diag -= dudx[a][a];
        diag *= mu * 2.0 / 3.0;
        for (int i = 0; i < D; i++) {
            for (int j = 0; j < D; j++) tauMC[i][j] = mu * dudx[j][i];
            tauMC[i][i] += diag;
        }
    }
    for (int i = 0; i < D; i++) divT[i] = dotD(D, tauMC[i], S);
    double dm = 0.0;
    for (int i = 0; i < D; i++) {
        delta[i] = d[i] * S_mag * S_mag / dotD(D, S, d);
        dm += delta[i] * delta[i];
        K[i] = S[i] - delta[i];
    }
    dm = sqrt(dm);
    for (int i = 0; i < D; i++) lapU[i] = mu * (dm * (aU[i] - cU[i]) * dmi + dotD(D, K, dudx[i]));
    double lapT = k_ * (dm * (aT - cT) * dmi + dotD(D, K, dTdx));
    for (int i = 0; i < D; i++) sU[i] = INTERP_LINEAR(w, cell->sigmaU[i], nb->sigmaU[i]);
    if (nb->is_ghost) {
        double Uf[3];
        for (int i = 0; i < D; i++) Uf[i] = INTERP_LINEAR(w, cU[i], aU[i]);
        for (int i = 0; i < D; i++) sU[i] = dotD(D, Uf, tau[i]);
    }
    double divS = dotD(D, sU, S);
    for (int i = 0; i < D; i++) rhs[i + 1] += divT[i] + lapU[i];
    rhs[D + 1] += divS + lapT;
}

static void convective_M1(const oracle *o, const cellgeo *G, int f, const solvars *cell, const solvars *nb,
                          double cri, double ari, const double *cU, const double *aU, double cR, double aR,
                          double *S_mag_out, double *rhs)
{
    const int D = o->D;
    const double w = G->w[f], omw = 1.0 - w;
    const double *S = G->S[f];
    double rhoPos = INTERP_LINEAR(w, cell->q[0], nb->q[0]);
    double rhoNeg = INTERP_LINEAR(omw, nb->q[0], cell->q[0]);
    double rPi = 1.0 / rhoPos, rNi = 1.0 / rhoNeg;
    double rUP[3], rUN[3], uP[3], uN[3];
    for (int k = 0; k < D; k++) {
        rUP[k] = INTERP_LINEAR(w, cell->q[k + 1], nb->q[k + 1]);
        rUN[k] = INTERP_LINEAR(omw, nb->q[k + 1], cell->q[k + 1]);
    }
    double ce = 2 * cell->q[D + 1] * cri;
    double ae = 2 * nb->q[D + 1] * ari;
    for (int k = 0; k < D; k++) { ce -= cU[k] * cU[k]; ae -= aU[k] * aU[k]; }
    ce *= 0.5;
    ae *= 0.5;
    double eP = INTERP_LINEAR(w, ce, ae), eN = INTERP_LINEAR(omw, ae, ce);
    double RP = INTERP_LINEAR(w, cR, aR), RN = INTERP_LINEAR(omw, aR, cR);
    double pP = rhoPos * RP, pN = rhoNeg * RN;
    double cP = sqrt(o->gamma * cR), cN = sqrt(o->gamma * aR);
    double cPos = INTERP_LINEAR(w, cP, cN), cNeg = INTERP_LINEAR(omw, cN, cP);
    double phP = 0.0, phN = 0.0;
    for (int k = 0; k < D; k++) {
        uP[k] = rUP[k] * rPi;
        uN[k] = rUN[k] * rNi;
        phP += uP[k] * S[k];
        phN += -uN[k] * S[k];
    }
    double Sm = 0.;
    for (int k = 0; k < D; k++) Sm += S[k] * S[k]; /* pow(S, 2) in the reference */
    Sm = sqrt(Sm);
    *S_mag_out = Sm;
    double a1_ = phP + cPos * Sm, a2_ = -phN + cNeg * Sm;
    double psiP = (a1_ < a2_) ? a2_ : a1_;  /* std::max({...}): first largest */
    if (psiP < 0.0) psiP = 0.0;
    double b1_ = phP - cPos * Sm, b2_ = -phN - cNeg * Sm;
    double psiN = (b2_ < b1_) ? b2_ : b1_;  /* std::min({...}): first smallest */
    if (0.0 < psiN) psiN = 0.0;
    double a0 = 1.0 / (psiP - psiN);
    double a1 = psiP * psiN;
    double aP = psiP * phP, aN = psiN * phN;
    double rEP = eP, rEN = eN;
    for (int k = 0; k < D; k++) { rEP += 0.5 * uP[k] * uP[k]; rEN += 0.5 * uN[k] * uN[k]; }
    rEP *= rhoPos;
    rEN *= rhoNeg;
    rhs[0] = -(aP * rhoPos + aN * rhoNeg + (rhoNeg - rhoPos) * a1) * a0;
    for (int k = 0; k < D; k++) {
        double phiUp = (aP * rUP[k] + aN * rUN[k] + (rUN[k] - rUP[k]) * a1) * a0 + (pP * psiP - pN * psiN) * a0 * S[k];
        rhs[k + 1] = -phiUp;
    }
    rhs[D + 1] = -(aP * rEP + aN * rEN + (rEN - rEP) * a1 + (aP * pP + aN * pN)) * a0;
}

static void convective_M2(const oracle *o, const cellgeo *G, int f, const solvars *cell, const solvars *nb,
                          double cri, double ari, const double *cU, const double *aU, double cR, double aR,
                          double *S_mag_out, double *rhs, int ausm)
{
    const int D = o->D;
    const double *S = G->S[f], *d = G->d[f];
    double rhoavg = 0.5 * (cell->q[0] + nb->q[0]);
    double rai = 1.0 / rhoavg;
    double rUa[3], ua[3];
    for (int k = 0; k < D; k++) rUa[k] = 0.5 * (cell->q[k + 1] + nb->q[k + 1]);
    double Ravg = 0.5 * (cR + aR);
    double pavg = rhoavg * Ravg;
    double cH = cell->q[D + 1] / cell->q[0] + cR;
    double aH = nb->q[D + 1] / nb->q[0] + aR;
    double Havg = 0.5 * (cH + aH);
    double cP = sqrt(o->gamma * cR), cN = sqrt(o->gamma * aR);
    double cavg = 0.5 * (cP + cN);
    (void)cri; (void)ari; (void)cU; (void)aU;
    if (ausm) {
        double cdiv = 0, ndiv = 0, ccurl = 0, ncurl = 0, cvm = 0, nvm = 0, unP = 0, unN = 0, nrm = 0;
        double cu[3], nu[3];
        for (int a = 0; a < D; a++) {
            cdiv += cell->dudx[a][a];
            ndiv += nb->dudx[a][a];
            for (int b = a + 1; b < D; b++) {
                ccurl += (cell->dudx[a][b] - cell->dudx[b][a]) * (cell->dudx[a][b] - cell->dudx[b][a]);
                ncurl += (nb->dudx[a][b] - nb->dudx[b][a]) * (nb->dudx[a][b] - nb->dudx[b][a]);
            }
            cu[a] = cell->q[a + 1] / cell->q[0];
            cvm += cu[a] * cu[a];
            nu[a] = nb->q[a + 1] / nb->q[0];
            nvm += nu[a] * nu[a];
            double uPos = interp_minmod(D, cu[a], nu[a], cell->U_grad[a], d, G->w[f], 1.0);
            double uNeg = interp_minmod(D, cu[a], nu[a], nb->U_grad[a], d, G->w[f], -1.0);
            unP += uPos * S[a];
            unN += uNeg * S[a];
            nrm += S[a] * S[a];
        }
        double cth = -(cdiv / sqrt(cdiv * cdiv + ccurl + 4e-2));
        double nth = -(ndiv / sqrt(ndiv * ndiv + ncurl + 4e-2));
        cth = (cth < 0.) ? 0. : cth;  /* std::max(x, 0.) */
        nth = (nth < 0.) ? 0. : nth;
        double tha = 0.5 * (cth + nth);
        double r = cell->q[0], E = cell->q[D + 1] / r;
        double cp_ = r * o->gm1 * (E - 0.5 * cvm);
        r = nb->q[0]; E = nb->q[D + 1] / r;
        double np_ = r * o->gm1 * (E - 0.5 * nvm);
        double rhoP = interp_minmod(D, cell->q[0], nb->q[0], cell->rho_grad, d, G->w[f], 1.0);
        double rhoN = interp_minmod(D, cell->q[0], nb->q[0], nb->rho_grad, d, G->w[f], -1.0);
        double pP = interp_minmod(D, cp_, np_, cell->p_grad, d, G->w[f], 1.0);
        double pN = interp_minmod(D, cp_, np_, nb->p_grad, d, G->w[f], -1.0);
        double Msq = (unP * unP + unN * unN) / (2. * cavg * cavg * nrm);
        double MP = unP / (sqrt(nrm) * cavg);
        double MN = unN / (sqrt(nrm) * cavg);
        double Minf = 0.2;
        double mx = (Msq < Minf * Minf) ? Minf * Minf : Msq;  /* std::max */
        double M0 = sqrt((mx < 1.0) ? mx : 1.0);                /* std::min(1.0, .) */
        double fa = M0 * (2.0 - M0);
        double alpha = 3.0 * (5.0 * fa * fa - 4.0) / 16.0;
        double phalf = pN * (p5Pos(MN, alpha) - p5Neg(MN, alpha)) - pP * (p5Pos(MP, alpha) - p5Neg(MP, alpha));
        double pu = -0.75 * p5Pos(MP, alpha) * p5Neg(MN, alpha) * (rhoP + rhoN) * cavg * fa * (unN - unP);
        pavg += tha * (pu - 0.5 * phalf);
    }
    double phi = 0.0;
    for (int k = 0; k < D; k++) {
        ua[k] = rUa[k] * rai;
        phi += ua[k] * S[k];
    }
    double Sm = 0.;
    for (int k = 0; k < D; k++) Sm += S[k] * S[k]; /* pow(S, 2) in the reference */
    Sm = sqrt(Sm);
    *S_mag_out = Sm;
    rhs[0] = -rhoavg * phi;
    for (int k = 0; k < D; k++) rhs[k + 1] = -(rUa[k] * phi + pavg * S[k]);
    rhs[D + 1] = -(rhoavg * Havg * phi);
}

static void one_rk_step(oracle *o, int rk, double dt)
{
    const int D = o->D;
    const int chk = (o->have_res && rk == 0);
    for (int c = 0; c < o->nc; c++) {
        cellgeo *G = &o->g[c];
        solvars *cell = &o->v[c];
        double cri = 1.0 / cell->q[0];
        double cU[3];
        for (int k = 0; k < D; k++) cU[k] = cell->q[k + 1] * cri;
        double cR = compute_Rpsi(o, cell->q);
        double cT = cR * o->R_inv;
        for (int f = 0; f < G->nvalid; f++) {
            solvars *nb = &o->v[G->neigh[f]];
            double ari = 1.0 / nb->q[0];
            double aU[3];
            for (int k = 0; k < D; k++) aU[k] = nb->q[k + 1] * ari;
            double aR = compute_Rpsi(o, nb->q);
            double aT = aR * o->R_inv;
            double rhs[5], Sm;
            if (o->solver == 0) convective_M1(o, G, f, cell, nb, cri, ari, cU, aU, cR, aR, &Sm, rhs);
            else convective_M2(o, G, f, cell, nb, cri, ari, cU, aU, cR, aR, &Sm, rhs, o->solver == 2);
            viscous_face(o, G, f, cell, nb, cU, aU, cT, aT, Sm, rhs);
            if (chk)
                for (int k = 0; k < D + 2; k++) { cell->RES[k] += rhs[k]; nb->RES[k] -= rhs[k]; }
            for (int k = 0; k < D + 2; k++) {
                cell->dq[k] += dt * rhs[k] * cell->vol_inv;
                nb->dq[k] -= dt * rhs[k] * nb->vol_inv;
            }
        }
        double sg = G->sigma;
        cell->dq[0] += dt * sg * (o->rho_inf - cell->q[0]);
        for (int k = 0; k < D; k++) cell->dq[k + 1] += dt * sg * (o->rho_inf * o->U_inf[k] - cell->q[k + 1]);
        cell->dq[D + 1] += dt * sg * (o->rho_inf * o->E_inf - cell->q[D + 1]);
        for (int k = 0; k < D + 2; k++) cell->q[k] += o->Bk[rk] * cell->dq[k];
    }
    if (chk) {
        /* per-submesh sums, submesh 0 first (RES accumulates across them) */
        for (int c = 0; c < o->nc; c++)
            for (int k = 0; k < D + 2; k++) o->RESsum[k] += o->v[c].RES[k] * o->v[c].RES[k];
    }
}

double lfmo_compute_cfl(void *h, double dt)
{
    oracle *o = h;
    const int D = o->D;
    double cmax = 0.;
    for (int c = 0; c < o->nc; c++) {
        double cfl = 0;
        double ri = 1.0 / o->v[c].q[0];
        for (int f = 0; f < o->g[c].nface; f++) {
            double fl = 0.;
            for (int k = 0; k < D; k++) fl += o->v[c].q[k + 1] * o->g[c].S[f][k];
            cfl += fabs(fl);
        }
        cfl *= ri * o->v[c].vol_inv * dt;
        if (cmax < cfl) cmax = cfl;
    }
    return cmax;
}

double lfmo_compute_dt(void *h, double max_co)
{
    oracle *o = h;
    const int D = o->D;
    double dmin = 100.0;
    for (int c = 0; c < o->nc; c++) {
        double sum = 0;
        double rho = o->v[c].q[0];
        for (int f = 0; f < o->g[c].nface; f++) {
            double sM = 0.;
            for (int k = 0; k < D; k++) sM += o->g[c].S[f][k] * o->g[c].S[f][k];
            sM = sqrt(sM);
            double fl = 0.;
            for (int k = 0; k < D; k++) fl += o->v[c].q[k + 1] * o->g[c].S[f][k] / sM;
            sum += fabs(fl);
        }
        double vol = 1.0 / o->v[c].vol_inv;
        double dt = max_co * rho * 2.0 * vol / sum;
        if (dmin > dt) dmin = dt;
    }
    return dmin;
}

static void run_stage(oracle *o, int rk)
{
    prepare_for_RKstep(o, rk);
    set_boundary_conditions(o);
    if (o->minmod) {
        if (o->solver == 2) calc_gradients_M2AUSM(o);
        else calc_gradients(o);
    }
    if (o->laminar) calc_VIS(o);
    else calc_VIS_Smagorinsky(o);
    one_rk_step(o, rk, o->dt);
}

int lfmo_stage_phase(void *h, int rk, int begin_step, int phase)
{
    oracle *o = h;
    if (phase == 0) {
        if (begin_step) {
            prepare_for_timestep(o);
            for (int k = 0; k < 5; k++) o->RESsum[k] = 0.0;
        }
        prepare_for_RKstep(o, rk);
        set_boundary_conditions(o);
        if (o->minmod) {
            if (o->solver == 2) calc_gradients_M2AUSM(o);
            else calc_gradients(o);
        }
        if (o->laminar) calc_VIS(o);
        else calc_VIS_Smagorinsky(o);
    } else {
        one_rk_step(o, rk, o->dt);
    }
    return 0;
}

void lfmo_set_dt(void *h, double dt) { ((oracle *)h)->dt = dt; }

/* ------------------------------------------------------------------ */
/* partitions: set-up messages and halo payloads                       */
/* ------------------------------------------------------------------ */
static int nb_index(const oracle *o, int rank)
{
    for (int b = 0; b < o->nnb; b++)
        if (o->nb_rank[b] == rank) return b;
    return -1;
}

int lfmo_halo_nbrs(void *h, int *ranks)
{
    oracle *o = h;
    if (ranks)
        for (int b = 0; b < o->nnb; b++) ranks[b] = o->nb_rank[b];
    return o->nnb;
}

/* mesh_reader.cpp:521-555: face id, owner cell, cell centre - face centre */
int lfmo_halo_export(void *h, int rank, int *keys, int *cells, double *disp)
{
    oracle *o = h;
    int b = nb_index(o, rank);
    if (b < 0) return -1;
    int n = o->nb_off[b + 1] - o->nb_off[b];
    for (int i = 0; i < n; i++) {
        int k = o->nb_off[b] + i;
        if (keys) keys[i] = o->pf_key[k];
        if (cells) cells[i] = o->pf_cell[k];
        if (disp)
            for (int q = 0; q < 3; q++) disp[3 * i + q] = o->pf_xc[3 * k + q] - o->pf_xf[3 * k + q];
    }
    return n;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* mesh_reader.cpp:577-605 (v_neigh from the received displacement),
 * cfd_v0.cpp:800-855 (d, linear weight), cfd_v0.cpp:616-700 (the
 * neighbour's send list = its cells on the shared faces, ascending) */
int lfmo_halo_import(void *h, int rank, int n, const int *keys, const int *cells, const double *disp)
{
    oracle *o = h;
    const int D = o->D;
    int b = nb_index(o, rank);
    if (b < 0 || n != o->nb_off[b + 1] - o->nb_off[b]) return -1;
    int *uniq = malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; i++) uniq[i] = cells[i];
    qsort(uniq, (size_t)n, sizeof(int), cmp_int);
    int nu = 0;
    for (int i = 0; i < n; i++)
        if (nu == 0 || uniq[nu - 1] != uniq[i]) uniq[nu++] = uniq[i];
    free(o->nb_recv[b]);
    o->nb_recv[b] = uniq;
    o->nb_nrecv[b] = nu;
    const int base = o->nc + o->ng + o->nb_off[b];
    for (int i = 0; i < nu; i++) {
        solvars *gv = &o->v[base + i];
        gv->sm_id = GHOST_SM;
        gv->id = uniq[i];
        gv->is_ghost = 0;
        gv->vol_inv = 0.0;
    }
    for (int k = o->nb_off[b]; k < o->nb_off[b + 1]; k++) {
        int i = 0;
        while (i < n && keys[i] != o->pf_key[k]) i++;
        if (i == n) return -2;
        int lo = 0, hi = nu - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (uniq[mid] < cells[i]) lo = mid + 1; else hi = mid;
        }
        cellgeo *G = &o->g[o->pf_cell[k]];
        int s = o->pf_slot[k];
        double dP[3], dN[3], w = 0.0, div = 0.0;
        for (int q = 0; q < D; q++) {
            G->d[s][q] = disp[3 * i + q] + o->pf_xf[3 * k + q] - o->pf_xc[3 * k + q];
            dP[q] = o->pf_xc[3 * k + q] - o->pf_xf[3 * k + q];
        }
        for (int q = 0; q < D; q++) dN[q] = dP[q] + G->d[s][q];
        for (int q = 0; q < D; q++) {
            w += G->S[s][q] * dN[q];
            div += (-G->S[s][q] * dP[q] + G->S[s][q] * dN[q]);
        }
        G->w[s] = w / div;
        G->neigh[s] = base + lo;
    }
    o->nb_imported[b] = 1;
    return 0;
}

/* one cell's record: q, dudx, dTdx, tauMC, sigmaU (cfd_v0.cpp:3303-3331) */
static int record(const oracle *o, solvars *v, int kind, double *buf, int dir)
{
    const int D = o->D;
    int m = 0;
#define XFER(x) do { if (dir) buf[m] = (x); else (x) = buf[m]; m++; } while (0)
    if (kind & 1)
        for (int i = 0; i < D + 2; i++) XFER(v->q[i]);
    if (kind & 2) {
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) XFER(v->dudx[i][j]);
        for (int i = 0; i < D; i++) XFER(v->dTdx[i]);
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) XFER(v->tauMC[i][j]);
        for (int i = 0; i < D; i++) XFER(v->sigmaU[i]);
    }
    if (kind & 4) {
        /* not a reference payload: the M2AUSM minmod gradients, which the
         * reference never exchanges (so its partitioned runs differ from
         * serial at processor faces); used to show that this is the only
         * difference */
        for (int i = 0; i < D; i++) XFER(v->rho_grad[i]);
        for (int i = 0; i < D; i++) XFER(v->p_grad[i]);
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) XFER(v->U_grad[i][j]);
    }
#undef XFER
    return m;
}

int lfmo_halo_pack(void *h, int rank, int kind, double *buf)
{
    oracle *o = h;
    int b = nb_index(o, rank);
    if (b < 0) return -1;
    /* send list: boundary-submesh cells with a face to this rank, in cell
     * order, each once (assign_pointers, cfd_v0.cpp:616-639) */
    int off = 0, last = -1;
    int n = o->nb_off[b + 1] - o->nb_off[b];
    int *cl = malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; i++) cl[i] = o->pf_cell[o->nb_off[b] + i];
    qsort(cl, (size_t)n, sizeof(int), cmp_int);
    for (int i = 0; i < n; i++) {
        if (cl[i] == last) continue;
        last = cl[i];
        off += record(o, &o->v[cl[i]], kind, buf + off, 1);
    }
    free(cl);
    return off;
}

/* unpack_mpi_data[_solvars|_viscous], cfd_v0.cpp:3608-3692 */
int lfmo_halo_unpack(void *h, int rank, int kind, const double *buf)
{
    oracle *o = h;
    int b = nb_index(o, rank);
    if (b < 0 || !o->nb_imported[b]) return -1;
    int off = 0;
    const int base = o->nc + o->ng + o->nb_off[b];
    for (int i = 0; i < o->nb_nrecv[b]; i++)
        off += record(o, &o->v[base + i], kind, (double *)buf + off, 0);
    return off;
}

int lfmo_stage(void *h, int rk, int begin_step)
{
    oracle *o = h;
    if (begin_step) {
        prepare_for_timestep(o);
        for (int k = 0; k < 5; k++) o->RESsum[k] = 0.0;
    }
    run_stage(o, rk);
    return 0;
}

int lfmo_run(void *h, int nsteps)
{
    oracle *o = h;
    for (int s = 0; s < nsteps; s++) {
        prepare_for_timestep(o);
        for (int k = 0; k < 5; k++) o->RESsum[k] = 0.0;
        for (int rk = 0; rk < o->nrk; rk++) run_stage(o, rk);
        if (o->adjust) o->dt = lfmo_compute_dt(o, o->max_co);
        else o->cfl = lfmo_compute_cfl(o, o->dt);
    }
    return 0;
}

double lfmo_get_dt(void *h) { return ((oracle *)h)->dt; }
double lfmo_get_cfl(void *h) { return ((oracle *)h)->cfl; }

int lfmo_get_res(void *h, double *res)
{
    oracle *o = h;
    for (int k = 0; k < o->D + 2; k++) res[k] = o->RESsum[k];
    return 0;
}

int lfmo_n_ghost(void *h) { return ((oracle *)h)->ng; }

int lfmo_ncomp(void *h, int which)
{
    int D = ((oracle *)h)->D;
    switch (which) {
    case LFMO_F_Q: case LFMO_F_DQ: case LFMO_F_RES: return D + 2;
    case LFMO_F_DUDX: case LFMO_F_TAUMC: case LFMO_F_UGRAD: case LFMO_F_RHOUGRAD: return D * D;
    case LFMO_F_DTDX: case LFMO_F_SIGMAU: case LFMO_F_RHOGRAD: case LFMO_F_PGRAD:
    case LFMO_F_RHOEGRAD: case LFMO_F_RPSIGRAD: case LFMO_F_CGRAD: return D;
    case LFMO_F_CENTRE: return 3;
    default: return 1;
    }
}

int lfmo_get_field(void *h, int which, double *out)
{
    oracle *o = h;
    const int D = o->D, nc = o->nc;
    int nk = lfmo_ncomp(h, which);
    for (int ci = 0; ci < nc; ci++) {
        int c = o->orig[ci];
        const solvars *v = &o->v[ci];
        for (int k = 0; k < nk; k++) {
            double x = 0;
            switch (which) {
            case LFMO_F_Q: x = v->q[k]; break;
            case LFMO_F_DQ: x = v->dq[k]; break;
            case LFMO_F_DUDX: x = v->dudx[k / D][k % D]; break;
            case LFMO_F_DTDX: x = v->dTdx[k]; break;
            case LFMO_F_TAUMC: x = v->tauMC[k / D][k % D]; break;
            case LFMO_F_SIGMAU: x = v->sigmaU[k]; break;
            case LFMO_F_RHOGRAD: x = v->rho_grad[k]; break;
            case LFMO_F_PGRAD: x = v->p_grad[k]; break;
            case LFMO_F_UGRAD: x = v->U_grad[k / D][k % D]; break;
            case LFMO_F_RES: x = v->RES[k]; break;
            case LFMO_F_VOLINV: x = v->vol_inv; break;
            case LFMO_F_SIGMA: x = o->g[ci].sigma; break;
            case LFMO_F_CENTRE: x = o->centre[3 * c + k]; break;
            case LFMO_F_VOLUME: x = o->volume[c]; break;
            case LFMO_F_SUBMESH: x = v->sm_id; break;
            case LFMO_F_RHOUGRAD: x = v->rhoU_grad[k / D][k % D]; break;
            case LFMO_F_RHOEGRAD: x = v->rhoE_grad[k]; break;
            case LFMO_F_RPSIGRAD: x = v->Rpsi_grad[k]; break;
            case LFMO_F_CGRAD: x = v->c_grad[k]; break;
            default: return -1;
            }
            out[(size_t)k * nc + c] = x;
        }
    }
    return 0;
}

int lfmo_set_field(void *h, int which, const double *in)
{
    oracle *o = h;
    const int nc = o->nc;
    if (which != LFMO_F_Q && which != LFMO_F_DQ) return -1;
    for (int ci = 0; ci < nc; ci++) {
        int c = o->orig[ci];
        for (int k = 0; k < o->D + 2; k++) {
            double x = in[(size_t)k * nc + c];
            if (which == LFMO_F_Q) o->v[ci].q[k] = x; else o->v[ci].dq[k] = x;
        }
    }
    return 0;
}

int lfmo_get_ghosts(void *h, double *q, int *cell, int *role)
{
    oracle *o = h;
    for (int g = 0; g < o->ng; g++) {
        if (q)
            for (int k = 0; k < o->D + 2; k++) q[(size_t)k * o->ng + g] = o->v[o->nc + g].q[k];
        if (cell) cell[g] = o->orig[o->ghost_cell[g]];
        if (role) role[g] = o->ghost_role[g];
    }
    return o->ng;
}

int lfmo_get_slots(void *h, int *nface, int *nvalid, int *nbr, double *S, double *d, double *w)
{
    oracle *o = h;
    for (int ci = 0; ci < o->nc; ci++) {
        int c = o->orig[ci];
        const cellgeo *G = &o->g[ci];
        nface[c] = G->nface;
        nvalid[c] = G->nvalid;
        for (int s = 0; s < MAXF; s++) {
            int j = c * MAXF + s;
            if (s < G->nface) {
                int n = G->neigh[s];
                nbr[j] = n < o->nc ? o->orig[n] : -1 - (n - o->nc);
                w[j] = G->w[s];
                for (int k = 0; k < 3; k++) { S[3 * j + k] = G->S[s][k]; d[3 * j + k] = G->d[s][k]; }
            } else {
                nbr[j] = -1000000;
                w[j] = 0;
                for (int k = 0; k < 3; k++) { S[3 * j + k] = 0; d[3 * j + k] = 0; }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* post-processing: averages and wall forces                           */
/* ------------------------------------------------------------------ */
int lfmo_set_boundary_conditions(void *h)
{
    set_boundary_conditions((oracle *)h);
    return 0;
}

/* postProcAverage(time_step), cfd_v0.cpp:3083-3103: for n > 0 add the cell
 * pressure to pAVG and (p - pAVG/n)^2 to pRMS, pAVG already holding p. */
int lfmo_average(void *h, int n)
{
    oracle *o = h;
    const int D = o->D;
    if (!o->pAVG) {
        o->pAVG = calloc((size_t)o->nc, sizeof(double));
        o->pRMS = calloc((size_t)o->nc, sizeof(double));
    }
    if (n <= 0) return 0;
    for (int ci = 0; ci < o->nc; ci++) {
        const double *q = o->v[ci].q;
        double r = q[0], velMag = 0.0;
        for (int k = 1; k <= D; k++) {
            double vel = q[k] / r;
            velMag += vel * vel;
        }
        double E = q[D + 1] / r;
        double p = (r * o->gm1) * (E - 0.5 * velMag);
        o->pAVG[ci] += p;
        o->pRMS[ci] += pow(p - o->pAVG[ci] / n, 2);
    }
    return 0;
}

/* updateAverageField(n), cfd_v0.cpp:2966-2970 with the field updates of
 * polyMeshReaderOF.cpp:792-861: pAvg = pAVG * (1/n), pRMS = sqrt(pRMS * (1/n)),
 * original cell order. */
int lfmo_average_fields(void *h, int n, double *pavg, double *prms)
{
    oracle *o = h;
    if (!o->pAVG || n <= 0) return -1;
    const double f = 1.0 / n;
    for (int ci = 0; ci < o->nc; ci++) {
        pavg[o->orig[ci]] = o->pAVG[ci] * f;
        prms[o->orig[ci]] = sqrt(o->pRMS[ci] * f);
    }
    return 0;
}

int lfmo_walls(void *h, int *patches)
{
    oracle *o = h;
    for (int i = 0; i < o->n_walls && patches; i++) patches[i] = o->wall_patch[i];
    return o->n_walls;
}

/* postProcForces, cfd_v0.cpp:3169-3233, per wall patch in m_nWallBCList
 * order: out[w*2D + k] = Fp_k, out[w*2D + D + k] = Fv_k.  Each wall face is
 * looked up as the reference does, by boundaries[patch][i] = {cell, slot}
 * with the slot its face had BEFORE reorder_faces, in the cell's arrays as
 * they are AFTER it (S, d and neighbour of whatever face moved there).
 * own_face != 0 uses the wall face's own S, d and ghost instead.  The
 * neighbour's state is what it holds now: for boundary ghosts, the fill of
 * the last set_boundary_conditions. */
int lfmo_forces(void *h, int own_face, double *out)
{
    oracle *o = h;
    const int D = o->D;
    const double mu = o->mu;
    for (int w = 0; w < o->n_walls; w++) {
        double Fpre[3] = {0.0, 0.0, 0.0}, Fvis[3] = {0.0, 0.0, 0.0};
        for (int g = 0; g < o->ng; g++) {
            if (o->ghost_patch[g] != o->wall_patch[w]) continue;
            const int ci = o->ghost_cell[g];
            const cellgeo *G = &o->g[ci];
            int face = o->ghost_slot[g];
            if (own_face)
                for (face = 0; G->neigh[face] != o->nc + g; face++) {}
            const solvars *vars = &o->v[ci];
            const solvars *neigh = &o->v[G->neigh[face]];
            double r = vars->q[0], Umag = 0.0;
            for (int k = 0; k < D; k++) {
                double u = vars->q[k + 1] / r;
                Umag += u * u;
            }
            Umag = sqrt(Umag);
            double E = vars->q[D + 1] / r;
            double p = (r * o->gm1) * (E - 0.5 * Umag * Umag);
            for (int k = 0; k < D; k++) Fpre[k] += G->S[face][k] * (p - o->p_inf);
            double dmag_inv = 1.0 / magD(D, G->d[face]);
            double dn[3], dudx[3][3], tau[3][3];
            for (int k = 0; k < D; k++) dn[k] = G->d[face][k] * dmag_inv;
            for (int i = 0; i < D; i++) {
                double cu = vars->q[i + 1] / vars->q[0];
                double au = neigh->q[i + 1] / neigh->q[0];
                for (int j = 0; j < D; j++) dudx[i][j] = (au - cu) * dn[j] * dmag_inv;
            }
            for (int a = 0; a < D; a++) {
                tau[a][a] = 2.0 * dudx[a][a];
                for (int a1 = a + 1; a1 < D + a; a1++) {
                    int a2 = a1 % D;
                    tau[a][a2] = mu * (dudx[a][a2] + dudx[a2][a]);
                    tau[a][a] -= dudx[a2][a2];
                }
                tau[a][a] *= 2.0 / 3.0 * mu;
            }
            for (int a = 0; a < D; a++)
                for (int b = 0; b < D; b++) Fvis[a] -= tau[a][b] * G->S[face][b];
        }
        for (int k = 0; k < D; k++) {
            out[w * 2 * D + k] = Fpre[k];
            out[w * 2 * D + D + k] = Fvis[k];
        }
    }
    return o->n_walls;
}

/* ------------------------------------------------------------------ */
/* unit-level entry points for the oracle's own closed-form tests      */
/* ------------------------------------------------------------------ */
double lfmo_minmod(int D, double cphi, double aphi, const double *grad, const double *d, double wl, double flux)
{
    return interp_minmod(D, cphi, aphi, grad, d, wl, flux);
}

void lfmo_p5(double M, double alpha, double *pos_neg)
{
    pos_neg[0] = p5Pos(M, alpha);
    pos_neg[1] = p5Neg(M, alpha);
}

/* rhs of one face (convective + viscous) as one_rk_step_* computes it, for
 * given cell states; vis/grad arrays hold dudx[D*D], dTdx[D], tauMC[D*D],
 * sigmaU[D], rho_grad[D], p_grad[D], U_grad[D*D] for P then N. */
int lfmo_face_rhs(const lfmo_params *prm, const double *qP, const double *qN, const double *S, const double *d,
                  double w, int ghost, const double *visP, const double *visN, double *rhs)
{
    oracle o;
    memset(&o, 0, sizeof o);
    const int D = prm->dim;
    o.D = D;
    o.solver = prm->solver;
    o.mu = prm->mu;
    o.cp = prm->cp;
    o.R = 8.31447 / prm->mol_weight * 1000;
    o.gamma = o.cp / (o.cp - o.R);
    o.gm1 = o.gamma - 1.0;
    o.R_inv = 1.0 / o.R;
    o.pr_inv = 1.0 / prm->pr;
    solvars c, n;
    memset(&c, 0, sizeof c);
    memset(&n, 0, sizeof n);
    cellgeo G;
    memset(&G, 0, sizeof G);
    for (int k = 0; k < D + 2; k++) { c.q[k] = qP[k]; n.q[k] = qN[k]; }
    for (int k = 0; k < D; k++) { G.S[0][k] = S[k]; G.d[0][k] = d[k]; }
    G.w[0] = w;
    n.is_ghost = ghost;
    solvars *cells[2] = {&c, &n};
    const double *vis[2] = {visP, visN};
    for (int s = 0; s < 2; s++) {
        const double *v = vis[s];
        solvars *x = cells[s];
        int o_ = 0;
        for (int i = 0; i < D; i++) for (int j = 0; j < D; j++) x->dudx[i][j] = v[o_++];
        for (int i = 0; i < D; i++) x->dTdx[i] = v[o_++];
        for (int i = 0; i < D; i++) for (int j = 0; j < D; j++) x->tauMC[i][j] = v[o_++];
        for (int i = 0; i < D; i++) x->sigmaU[i] = v[o_++];
        for (int i = 0; i < D; i++) x->rho_grad[i] = v[o_++];
        for (int i = 0; i < D; i++) x->p_grad[i] = v[o_++];
        for (int i = 0; i < D; i++) for (int j = 0; j < D; j++) x->U_grad[i][j] = v[o_++];
    }
    double cri = 1.0 / c.q[0], ari = 1.0 / n.q[0], cU[3], aU[3];
    for (int k = 0; k < D; k++) { cU[k] = c.q[k + 1] * cri; aU[k] = n.q[k + 1] * ari; }
    double cR = compute_Rpsi(&o, c.q), aR = compute_Rpsi(&o, n.q);
    double Sm;
    if (o.solver == 0) convective_M1(&o, &G, 0, &c, &n, cri, ari, cU, aU, cR, aR, &Sm, rhs);
    else convective_M2(&o, &G, 0, &c, &n, cri, ari, cU, aU, cR, aR, &Sm, rhs, o.solver == 2);
    viscous_face(&o, &G, 0, &c, &n, cU, aU, cR * o.R_inv, aR * o.R_inv, Sm, rhs);
    return 0;
}
