// kernels_fp64.hpp -- gfx950 kernels of the plain FP64 time step.
//
// One RK stage is four launches over SoA arrays:
//   bc_fill      one thread per boundary ghost   (set_boundary_conditions)
//   grad_vis     one thread per cell, gathering its faces in the reference's
//                processing order (calc_gradients[_M2AUSM] + calc_VIS)
//   face_flux    one thread per face, each face once (one_rk_step_* body)
//   cell_update  one thread per cell, gathering face fluxes in the
//                reference's accumulation order, then sponge and RK update
// The gathers replace the reference's owner/neighbour scatter: every cell
// sums the same terms in the same order as the reference's in-place loop,
// so no atomics are needed and the result does not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mesh_builder.hpp"

namespace lfm {

struct KArgs {
    int nc, nsm0, ng, nf, ntot;
    int ng_bnd;  // boundary ghosts [nc, nc+ng_bnd); halo ghosts after them
    const int32_t *fP, *fN;
    const double *S;                // [nf][D]: a face's area vector as one record
    const double *d, *w;            // d: [D][nf]
    const int32_t *g_off, *g_ent;   // gather lists
    const int32_t *s_off, *s_ent;   // slot lists
    const int32_t *ghost_cell, *ghost_role;
    const double *vinv, *sigma;
    const double *smag;             // [nc] Smagorinsky constants, nullptr when laminar
    double *q;                      // cell records (q64_index), ghosts included
    double *dq;                     // [(D+2)][nc]
    double *res_cell;               // [(D+2)][nc]
    // per-cell gradient records (GradRec64; ghosts' stay zero): dudx points
    // at the records, dTdx / rho_grad / p_grad / U_grad into them (component
    // stride 1, cell stride gstride); tauMC / sigmaU (partitions only) are
    // [component][ntot]
    double *dudx, *dTdx, *tauMC, *sigmaU;
    double *rho_grad, *p_grad, *U_grad;
    int gstride;
    double *frhs;                   // face records: [nf][4], then (3D) [nf]
    const double *dt;               // device scalar (deltaT)
    double *partial;                // reduction partials
    // gas
    double gamma, gm1, R_inv, mu, cp, pr_inv;
    double p_inf, T_inf, rho_inf, E_inf, U_inf[3];
};

#define INTERP_LINEAR(w, a, b) ((w) * (a) + (1.0 - (w)) * (b))

// MI355X dispatches workgroup b to XCD b % 8, and each XCD has its own L2.
// Remap so that every XCD runs one contiguous slice of the index space: the
// neighbours a face or cell gathers are then mostly in that XCD's L2.
// Bijection of [0, nb).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % kXcds, j = b / kXcds;
    const int per = nb / kXcds, rem = nb % kXcds;
    return x * per + (x < rem ? x : rem) + j;
}
#define LFM_TID(blk) (xcd_block((int)(blk), (int)gridDim.x) * (int)blockDim.x + (int)threadIdx.x)
#define LFM_BLK() xcd_block((int)blockIdx.x, (int)gridDim.x)

template <int D>
__device__ __forceinline__ double rpsi(const KArgs &a, const double *q) {
    const double ri = 1.0 / q[0];
    double s = q[1] * q[1];
#pragma unroll
    for (int k = 1; k < D; k++) s += q[k + 1] * q[k + 1];
    return a.gm1 * (q[D + 1] - 0.5 * s * ri) * ri;
}

template <int D>
__device__ __forceinline__ double dotD(const double *x, const double *y) {
    double s = x[0] * y[0];
#pragma unroll
    for (int k = 1; k < D; k++) s += x[k] * y[k];
    return s;
}

// Faces per cell whose gathers are issued together before any arithmetic
// (the cell's faces beyond it are gathered one at a time): every cell of a
// 2D quad/triangle mesh and of a 3D hex mesh fits.
template <int D>
struct GatherUnroll {
    static constexpr int value = D == 2 ? 4 : 6;
};

// FP64 loads and stores of data a kernel touches once, non-temporal where
// kStreamNT64 allows (levels as on the FP32 path: 2 the RK register, 3 the
// new states, the face kernel's face lists and geometry, the face fluxes,
// the gradient records, 1/V and sponge), so that they do not displace the
// gathered neighbour records in L2.  Every level in 2D and 3D (FP64 arrays
// outgrow the Infinity Cache sooner): 256^3 box 11.73 -> 11.38 ms per step,
// 4.2 M-cell O-grid 5.96 -> 5.65, 1 M-cell annulus 0.880 -> 0.848.
template <int D>
constexpr int kStreamNT64 = 3;
template <int D, int LEVEL, class T>
__device__ __forceinline__ T ld64_once(const T *p) {
    if constexpr (kStreamNT64<D> >= LEVEL) return __builtin_nontemporal_load(p);
    return *p;
}
template <int D, int LEVEL, class T>
__device__ __forceinline__ void st64_once(T *p, T v) {
    if constexpr (kStreamNT64<D> >= LEVEL)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
using d2v = double __attribute__((ext_vector_type(2)));
template <int D, int LEVEL>
__device__ __forceinline__ void st64x2_once(double *p, double x, double y) {
    const d2v v = {x, y};
    st64_once<D, LEVEL>(reinterpret_cast<d2v *>(p), v);
}

// face f's area vector (one 16-byte load in 2D)
template <int D>
__device__ __forceinline__ void load_S64(const double *S, int f, double (&s)[D]) {
    if constexpr (D == 2) {
        const double2 v = reinterpret_cast<const double2 *>(S)[f];
        s[0] = v.x;
        s[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < D; k++) s[k] = S[(size_t)f * D + k];
    }
}

// State layout: components 0-3 of a cell (or ghost) i as one 32-byte
// record, q[4 i + k], then component 4 (the 3D energy) as q[4 ntot + i]: a
// gathered neighbour state is two 16-byte loads (three in 3D).
__host__ __device__ __forceinline__ size_t q64_index(int k, size_t i, size_t ntot) {
    return k < 4 ? 4 * i + (size_t)k : 4 * ntot + i;
}

template <int D>
__device__ __forceinline__ void load_q(const KArgs &a, int i, double *q) {
    const double2 *p = reinterpret_cast<const double2 *>(a.q + 4 * (size_t)i);
    const double2 v0 = p[0], v1 = p[1];
    q[0] = v0.x;
    q[1] = v0.y;
    q[2] = v1.x;
    q[3] = v1.y;
    if (D == 3) q[4] = a.q[(size_t)4 * a.ntot + i];
}

template <int D>
__device__ __forceinline__ void store_q(double *qv, size_t ntot, int i, const double *q) {
    st64x2_once<D, 3>(qv + 4 * (size_t)i, q[0], q[1]);
    st64x2_once<D, 3>(qv + 4 * (size_t)i + 2, q[2], q[3]);
    if (D == 3) st64_once<D, 3>(qv + 4 * ntot + i, q[4]);
}

// One cell's gradients as one record, padded to whole 16-byte words: dudx
// (D*D), dTdx (D), then with RG (the M2AUSM scheme) the rho and p gradients
// (D each) and U_grad (D*D).
template <int D, bool RG>
struct GradRec64 {
    static constexpr int kDT = D * D;
    static constexpr int kRG = kDT + D;
    static constexpr int kPG = kRG + D;
    static constexpr int kUG = kPG + D;
    static constexpr int kN = RG ? kUG + D * D : kRG;
    static constexpr int kStride = (kN + 1) / 2 * 2;
};
inline int grad_record_stride64(int D, bool rg) {
    const int n = D * D + D + (rg ? 2 * D + D * D : 0);
    return (n + 1) / 2 * 2;
}

// face flux record: components 0-3 as 32 bytes per face, component 4 after
template <int D>
__device__ __forceinline__ void load_frhs(const KArgs &a, int f, double (&r)[D + 2]) {
    const double2 *p = reinterpret_cast<const double2 *>(a.frhs + 4 * (size_t)f);
    const double2 v0 = p[0], v1 = p[1];
    r[0] = v0.x;
    r[1] = v0.y;
    r[2] = v1.x;
    r[3] = v1.y;
    if (D == 3) r[4] = a.frhs[(size_t)4 * a.nf + f];
}

template <int D, bool RG>
__device__ __forceinline__ void load_grec(const KArgs &a, int c, double (&r)[GradRec64<D, RG>::kStride]) {
    using R = GradRec64<D, RG>;
    const double2 *p = reinterpret_cast<const double2 *>(a.dudx + (size_t)c * R::kStride);
#pragma unroll
    for (int k = 0; k < R::kStride / 2; k++) {
        const double2 v = p[k];
        r[2 * k] = v.x;
        r[2 * k + 1] = v.y;
    }
}

__device__ __forceinline__ double sgn(double x) { return x < 0.0 ? -1.0 : 1.0; }

template <int D>
__device__ __forceinline__ double interp_minmod(double cphi, double aphi, const double *grad, const double *d,
                                                double wl, double flux) {
    const double gradf = aphi - cphi + 1.0e-30;
    double gradcf = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) gradcf += d[k] * grad[k];
    double r;
    if (fabs(gradcf) >= 1000.0 * fabs(gradf))
        r = 2.0 * 1000.0 * sgn(gradcf) * sgn(gradf) - 1.0;
    else
        r = 2.0 * (gradcf / gradf) - 1.0;
    // std::max(std::min(r, 1), 0) of api/cfdv0_solver.h:364
    const double m1 = (1.0 < r) ? 1.0 : r;
    const double lim = (m1 < 0.0) ? 0.0 : m1;
    const double wt = lim * wl + (1.0 - lim) * (1.0 + flux) * 0.5;
    return wt * cphi + (1.0 - wt) * aphi;
}

__device__ __forceinline__ double p5pos(double M, double alpha) {
    const double M2Pos = 0.25 * (M + 1.0) * (M + 1.0);
    const double M2Neg = -0.25 * (M - 1.0) * (M - 1.0);
    const double M1Pos = 0.5 * (M + fabs(M));
    if (fabs(M) < 1) return M2Pos * ((2.0 - M) - 16.0 * alpha * M * M2Neg);
    return M1Pos / M;
}

__device__ __forceinline__ double p5neg(double M, double alpha) {
    const double M2Pos = 0.25 * (M + 1.0) * (M + 1.0);
    const double M2Neg = -0.25 * (M - 1.0) * (M - 1.0);
    const double M1Neg = 0.5 * (M - fabs(M));
    if (fabs(M) < 1) return M2Neg * ((-2.0 - M) + 16.0 * alpha * M * M2Pos);
    return M1Neg / M;
}

// ---------------------------------------------------------------------------
// Ghost state of one boundary face from its interior cell's state qi
// (cfd_v0.cpp:1010-1133); returns false for an unset patch (the ghost keeps
// q = 0).
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ bool ghost_state(const KArgs &a, int role, const double *qi, double *qg) {
    if (role == 1) {  // wall
        qg[0] = qi[0];
#pragma unroll
        for (int k = 0; k < D; k++) qg[k + 1] = -qi[k + 1];
        qg[D + 1] = qi[D + 1];
        return true;
    }
    if (role != 2 && role != 3) return false;
    const double rhoInt = qi[0];
    double Uint[D], u2i = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        Uint[k] = qi[k + 1] / rhoInt;
        u2i += Uint[k] * Uint[k];
    }
    const double pInt = (qi[D + 1] - 0.5 * u2i * rhoInt) * a.gm1;
    double rhoExt, pExt, Uext[D], u2e = 0.0;
    if (role == 2) {  // inlet
        const double TInt = pInt * a.R_inv / rhoInt;
        const double TExt = a.T_inf;
        rhoExt = rhoInt * TInt / TExt;
#pragma unroll
        for (int k = 0; k < D; k++) {
            Uext[k] = a.U_inf[k];
            u2e += Uext[k] * Uext[k];
        }
        pExt = pInt;
    } else {  // outlet
        pExt = a.p_inf;
        rhoExt = rhoInt * pExt / pInt;
#pragma unroll
        for (int k = 0; k < D; k++) {
            Uext[k] = Uint[k];
            u2e += Uext[k] * Uext[k];
        }
    }
    const double EExt = pExt / (rhoExt * a.gm1) + 0.5 * u2e;
    qg[0] = rhoExt;
#pragma unroll
    for (int k = 0; k < D; k++) qg[k + 1] = rhoExt * Uext[k];
    qg[D + 1] = rhoExt * EExt;
    return true;
}

// ---------------------------------------------------------------------------
// set_boundary_conditions (cfd_v0.cpp:1010-1133): ghost state per boundary
// face from its interior cell.  Ghost gradient copies are not stored: the
// interior gradients they copy are zero at this point of every stage, and
// the flux kernels read zero ghost gradients (see grad arrays' ghost part).
// In the fused step the update kernel refreshes the ghosts of the cells it
// updates, so this kernel only runs after the state is set from outside.
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) bc_fill(KArgs a) {
    const int g = LFM_TID(blockIdx.x);
    if (g >= a.ng) return;
    double qi[D + 2], qg[D + 2];
    load_q<D>(a, a.ghost_cell[g], qi);
    if (!ghost_state<D>(a, a.ghost_role[g], qi, qg)) return;
    const int gi = a.nc + g;
    store_q<D>(a.q, a.ntot, gi, qg);
}

// ---------------------------------------------------------------------------
// Green-Gauss gradients + viscous tensors, per cell (cfd_v0.cpp:1384-1567,
// 1744-1860).  The reference refreshes tau/sigmaU/tauMC after every touch of
// a cell; the last touch of a cell follows its last contribution, so the
// stored values equal the refresh of the completed sums done here.
// GRAD: 0 none, 1 calc_gradients (results unused by solvers 0/1: skipped),
// 2 calc_gradients_M2AUSM (rho, p, U gradients consumed by the M2AUSM flux).
// VISC false (mu = 0): dTdx, tau, sigmaU and tauMC only ever enter the flux
// multiplied by mu, so they are not formed; dudx is kept for the M2AUSM
// sensor.  (For solvers 0/1 with mu = 0 the kernel is not launched.)
// ---------------------------------------------------------------------------
// tauMC and sigmaU of one cell from its completed dudx (cfd_v0.cpp:1813-1834,
// Smagorinsky 1668-1684 when smag is given): the values calc_VIS leaves in
// the cell.  Used where they are stored and where a face recomputes them.
template <int D>
__device__ __forceinline__ void vis_tensors(const KArgs &a, const double (&dudx)[D][D], const double *qc,
                                            const double *smag_p, double (&tmc)[D][D], double (&sU)[D]) {
    double U[D];
#pragma unroll
    for (int k = 0; k < D; k++) U[k] = qc[k + 1] / qc[0];
    const double mu = a.mu;
    double tau[D][D];
#pragma unroll
    for (int p = 0; p < D; p++) {
        tau[p][p] = 2.0 * dudx[p][p];
#pragma unroll
        for (int b1 = p + 1; b1 < D + p; b1++) {
            const int b = b1 % D;
            tau[p][b] = mu * (dudx[p][b] + dudx[b][p]);
            tau[p][p] -= dudx[b][b];
        }
        tau[p][p] *= 2.0 / 3.0 * mu;
    }
    double diag = 0;
#pragma unroll
    for (int p = 0; p < D; p++) diag -= dudx[p][p];
    diag *= mu * 2.0 / 3.0;
#pragma unroll
    for (int i = 0; i < D; i++) {
#pragma unroll
        for (int j = 0; j < D; j++) tmc[i][j] = mu * dudx[j][i];
        tmc[i][i] += diag;
    }
    if (smag_p) {
        // calc_VIS_Smagorinsky: minus the eddy stress smag |S| S_ij, plus the
        // compressible correction smag |S| div u / 3
        const double smag = *smag_p;
        double strain = 0.0, divu = 0.0;
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                const double sij = 0.5 * (dudx[i][j] + dudx[j][i]);
                strain += sij * sij;
            }
            divu += dudx[i][i];
        }
        strain = sqrt(2.0 * strain);
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                const double sij = 0.5 * (dudx[i][j] + dudx[j][i]);
                tmc[i][j] -= smag * strain * sij;
            }
            tmc[i][i] += smag * strain * divu / 3.0;
        }
    }
#pragma unroll
    for (int i = 0; i < D; i++) sU[i] = dotD<D>(U, tau[i]);
}

template <int D, int GRAD, bool VISC, bool RG>
__global__ void __launch_bounds__(256) grad_vis(KArgs a) {
    const int c = LFM_TID(blockIdx.x);
    using R = GradRec64<D, RG>;
    // every thread reaches the record write-out below (LDS staging)
    const bool valid = c < a.nc;
    double rec[R::kStride];
#pragma unroll
    for (int k = 0; k < R::kStride; k++) rec[k] = 0.0;
    if (valid) {
    const double vinv = a.vinv[c];
    double dudx[D][D], dTdx[D];
    double rg[D], pg[D], ug[D][D];
#pragma unroll
    for (int i = 0; i < D; i++) {
        dTdx[i] = 0.0;
        rg[i] = 0.0;
        pg[i] = 0.0;
#pragma unroll
        for (int j = 0; j < D; j++) {
            dudx[i][j] = 0.0;
            ug[i][j] = 0.0;
        }
    }
    double qself[D + 2];
    load_q<D>(a, c, qself);
    const double Rself = rpsi<D>(a, qself), riself = 1.0 / qself[0];
    // one face's contribution, in the face's own P/N order
    auto add = [&](const bool isP, const double w, const double (&S)[D], const double (&qo)[D + 2]) {
        double qP[D + 2], qN[D + 2];
        const double Ro = rpsi<D>(a, qo), rio = 1.0 / qo[0];
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            qP[k] = isP ? qself[k] : qo[k];
            qN[k] = isP ? qo[k] : qself[k];
        }
        const double cR = isP ? Rself : Ro, aR = isP ? Ro : Rself;
        const double cri = isP ? riself : rio, ari = isP ? rio : riself;
        double sv[D];
#pragma unroll
        for (int k = 0; k < D; k++) sv[k] = isP ? S[k] * vinv : -S[k] * vinv;
        if (GRAD == 2) {
            double UU[D];
#pragma unroll
            for (int k = 0; k < D; k++) UU[k] = INTERP_LINEAR(w, qP[k + 1] / qP[0], qN[k + 1] / qN[0]);
            const double rho = INTERP_LINEAR(w, qP[0], qN[0]);
            double r = qP[0], E = qP[D + 1] / r, vm = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) vm += (qP[k + 1] / r) * (qP[k + 1] / r);
            const double cp_ = r * a.gm1 * (E - 0.5 * vm);
            r = qN[0];
            E = qN[D + 1] / r;
            vm = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) vm += (qN[k + 1] / r) * (qN[k + 1] / r);
            const double np_ = r * a.gm1 * (E - 0.5 * vm);
            const double pf = INTERP_LINEAR(w, cp_, np_);
#pragma unroll
            for (int i = 0; i < D; i++) {
                pg[i] += pf * sv[i];
#pragma unroll
                for (int j = 0; j < D; j++) ug[j][i] += UU[j] * sv[i];
                rg[i] += rho * sv[i];
            }
        }
        double fU[D];
#pragma unroll
        for (int k = 0; k < D; k++) fU[k] = INTERP_LINEAR(w, qP[k + 1] * cri, qN[k + 1] * ari);
        const double fT = VISC ? INTERP_LINEAR(w, cR, aR) * a.R_inv : 0.0;
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) dudx[i][j] += fU[i] * sv[j];
            if (VISC) dTdx[i] += fT * sv[i];
        }
    };
    const int e0 = a.g_off[c], ne = a.g_off[c + 1] - e0;
    // the first faces' gathers issued together before the sums (a short list
    // repeats its last face, whose terms are then not added)
    constexpr int U = GatherUnroll<D>::value;
    bool isPj[U];
    double wj[U], Sj[U][D], qj[U][D + 2];
#pragma unroll
    for (int j = 0; j < U; j++) {
        const int ent = a.g_ent[e0 + (j < ne ? j : ne - 1)];
        isPj[j] = ent >= 0;
        const int f = isPj[j] ? (ent & kFaceMask) : ~ent;
        wj[j] = a.w[f];
        load_S64<D>(a.S, f, Sj[j]);
        load_q<D>(a, isPj[j] ? a.fN[f] : a.fP[f], qj[j]);
    }
#pragma unroll
    for (int j = 0; j < U; j++)
        if (j < ne) add(isPj[j], wj[j], Sj[j], qj[j]);
    for (int e = e0 + U; e < e0 + ne; e++) {
        const int ent = a.g_ent[e];
        const bool p = ent >= 0;
        const int f = p ? (ent & kFaceMask) : ~ent;
        double Sf[D], q[D + 2];
        load_S64<D>(a.S, f, Sf);
        load_q<D>(a, p ? a.fN[f] : a.fP[f], q);
        add(p, a.w[f], Sf, q);
    }
    // refresh of tau / sigmaU / tauMC from the completed dudx (cfd_v0.cpp:1813-1834);
    // stored only where a halo payload needs them (otherwise each face
    // recomputes them from dudx)
    const size_t nt = a.ntot;
    if (VISC && a.tauMC) {
        double tmc[D][D], sU[D];
        vis_tensors<D>(a, dudx, qself, a.smag ? a.smag + c : nullptr, tmc, sU);
#pragma unroll
        for (int i = 0; i < D; i++) {
            a.sigmaU[i * nt + c] = sU[i];
#pragma unroll
            for (int j = 0; j < D; j++) a.tauMC[(i * D + j) * nt + c] = tmc[i][j];
        }
    }
#pragma unroll
    for (int i = 0; i < D; i++) {
#pragma unroll
        for (int j = 0; j < D; j++) rec[i * D + j] = dudx[i][j];
        rec[R::kDT + i] = VISC ? dTdx[i] : 0.0;
        if (RG && GRAD == 2) {
            rec[R::kRG + i] = rg[i];
            rec[R::kPG + i] = pg[i];
#pragma unroll
            for (int j = 0; j < D; j++) rec[R::kUG + i * D + j] = ug[i][j];
        }
    }
    if (!VISC) {
        // dTdx not formed: the formed components only (the record keeps
        // whatever a full calc_VIS left in the dTdx slots)
        double *out = a.dudx + (size_t)c * R::kStride;
#pragma unroll
        for (int k = 0; k < R::kN; k++)
            if (k < R::kDT || k >= R::kRG) out[k] = rec[k];
    }
    }
    if (VISC) {
        // the block's records are contiguous from its first cell: staged in
        // LDS, then written as consecutive 16-byte words (without minmod the
        // rho / p / U gradient slots are written zero, as they stay)
        constexpr int W = R::kStride / 2;
        __shared__ double2 st[256 * W];
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < W; k++) st[t * W + k] = make_double2(rec[2 * k], rec[2 * k + 1]);
        __syncthreads();
        const int c0 = c - t;
        const int nrec = a.nc - c0 < (int)blockDim.x ? a.nc - c0 : (int)blockDim.x;
        d2v *dst = reinterpret_cast<d2v *>(a.dudx + (size_t)c0 * R::kStride);
        for (int j = t; j < nrec * W; j += blockDim.x) st64_once<D, 3>(dst + j, d2v{st[j].x, st[j].y});
    }
}

// ---------------------------------------------------------------------------
// Face fluxes: convective (M1 2589-2675, M2 1968-2023, M2AUSM 2256-2366) plus
// the viscous terms shared by all three (2682-2790).  One thread per face.
// VISC false (mu = 0): every viscous term is mu or k = cp mu / Pr times a
// finite value, i.e. a zero the reference adds; it is not formed.
// STORED: tauMC / sigmaU read from the arrays grad_vis filled (partitions,
// whose halo payload carries them); otherwise formed here from dudx.
// In 2D, 3 waves per SIMD: the viscous M2AUSM face then fits 168 VGPRs
// without spills instead of 170 (2 waves), 82 -> 73 us per launch on 1 M
// cells; in 3D the same bound spills and is slower.
// ---------------------------------------------------------------------------
template <int D, int SOLVER, bool VISC, bool STORED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(D == 2 ? 3 : 2))) face_flux(KArgs a) {
    const int f = LFM_TID(blockIdx.x);
    if (f >= a.nf) return;
    const int P = ld64_once<D, 3>(a.fP + f), N = ld64_once<D, 3>(a.fN + f);
    // is_ghost: boundary ghosts only; halo ghosts carry the neighbour's tensors
    const bool ghost = (unsigned)(N - a.nc) < (unsigned)a.ng_bnd;
    const double w = ld64_once<D, 3>(a.w + f);
    const size_t nt = a.ntot;
    double S[D], d[D], qP[D + 2], qN[D + 2];
#pragma unroll
    for (int k = 0; k < D; k++) S[k] = ld64_once<D, 3>(a.S + (size_t)f * D + k);
#pragma unroll
    for (int k = 0; k < D; k++) d[k] = ld64_once<D, 3>(a.d + (size_t)k * a.nf + f);
    load_q<D>(a, P, qP);
    load_q<D>(a, N, qN);
    using R = GradRec64<D, SOLVER == 2>;
    double gr[2][R::kStride];  // gradient records of P and N
    if (VISC || SOLVER == 2) {
        load_grec<D, SOLVER == 2>(a, P, gr[0]);
        load_grec<D, SOLVER == 2>(a, N, gr[1]);
    }
    const double cri = 1.0 / qP[0], ari = 1.0 / qN[0];
    double cU[D], aU[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        cU[k] = qP[k + 1] * cri;
        aU[k] = qN[k + 1] * ari;
    }
    const double cR = rpsi<D>(a, qP), aR = rpsi<D>(a, qN);
    const double cT = cR * a.R_inv, aT = aR * a.R_inv;
    double rhs[D + 2];
    double Sm = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) Sm += S[k] * S[k];
    Sm = sqrt(Sm);

    if (SOLVER == 0) {
        const double omw = 1.0 - w;
        const double rhoPos = INTERP_LINEAR(w, qP[0], qN[0]);
        const double rhoNeg = INTERP_LINEAR(omw, qN[0], qP[0]);
        const double rPi = 1.0 / rhoPos, rNi = 1.0 / rhoNeg;
        double rUP[D], rUN[D], uP[D], uN[D];
#pragma unroll
        for (int k = 0; k < D; k++) {
            rUP[k] = INTERP_LINEAR(w, qP[k + 1], qN[k + 1]);
            rUN[k] = INTERP_LINEAR(omw, qN[k + 1], qP[k + 1]);
        }
        double ce = 2 * qP[D + 1] * cri;
        double ae = 2 * qN[D + 1] * ari;
#pragma unroll
        for (int k = 0; k < D; k++) {
            ce -= cU[k] * cU[k];
            ae -= aU[k] * aU[k];
        }
        ce *= 0.5;
        ae *= 0.5;
        const double eP = INTERP_LINEAR(w, ce, ae), eN = INTERP_LINEAR(omw, ae, ce);
        const double RP = INTERP_LINEAR(w, cR, aR), RN = INTERP_LINEAR(omw, aR, cR);
        const double pP = rhoPos * RP, pN = rhoNeg * RN;
        const double cP = sqrt(a.gamma * cR), cN = sqrt(a.gamma * aR);
        const double cPos = INTERP_LINEAR(w, cP, cN), cNeg = INTERP_LINEAR(omw, cN, cP);
        double phP = 0.0, phN = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            uP[k] = rUP[k] * rPi;
            uN[k] = rUN[k] * rNi;
            phP += uP[k] * S[k];
            phN += -uN[k] * S[k];
        }
        // std::max({x1, x2, 0}) / std::min({y1, y2, 0}): first extreme wins
        const double x1 = phP + cPos * Sm, x2 = -phN + cNeg * Sm;
        double psiP = (x1 < x2) ? x2 : x1;
        if (psiP < 0.0) psiP = 0.0;
        const double y1 = phP - cPos * Sm, y2 = -phN - cNeg * Sm;
        double psiN = (y2 < y1) ? y2 : y1;
        if (0.0 < psiN) psiN = 0.0;
        const double a0 = 1.0 / (psiP - psiN);
        const double a1 = psiP * psiN;
        const double aP = psiP * phP, aN = psiN * phN;
        double rEP = eP, rEN = eN;
#pragma unroll
        for (int k = 0; k < D; k++) {
            rEP += 0.5 * uP[k] * uP[k];
            rEN += 0.5 * uN[k] * uN[k];
        }
        rEP *= rhoPos;
        rEN *= rhoNeg;
        rhs[0] = -(aP * rhoPos + aN * rhoNeg + (rhoNeg - rhoPos) * a1) * a0;
#pragma unroll
        for (int k = 0; k < D; k++)
            rhs[k + 1] = -((aP * rUP[k] + aN * rUN[k] + (rUN[k] - rUP[k]) * a1) * a0 +
                           (pP * psiP - pN * psiN) * a0 * S[k]);
        rhs[D + 1] = -(aP * rEP + aN * rEN + (rEN - rEP) * a1 + (aP * pP + aN * pN)) * a0;
    } else {
        const double rhoavg = 0.5 * (qP[0] + qN[0]);
        const double rai = 1.0 / rhoavg;
        double rUa[D];
#pragma unroll
        for (int k = 0; k < D; k++) rUa[k] = 0.5 * (qP[k + 1] + qN[k + 1]);
        const double Ravg = 0.5 * (cR + aR);
        double pavg = rhoavg * Ravg;
        const double cH = qP[D + 1] / qP[0] + cR;
        const double aH = qN[D + 1] / qN[0] + aR;
        const double Havg = 0.5 * (cH + aH);
        if (SOLVER == 2) {
            const double cP = sqrt(a.gamma * cR), cN = sqrt(a.gamma * aR);
            const double cavg = 0.5 * (cP + cN);
            double cdiv = 0, ndiv = 0, ccurl = 0, ncurl = 0, cvm = 0, nvm = 0, unP = 0, unN = 0, nrm = 0;
            double cdu[D][D], ndu[D][D];
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) {
                    cdu[i][j] = gr[0][i * D + j];
                    ndu[i][j] = gr[1][i * D + j];
                }
#pragma unroll
            for (int p = 0; p < D; p++) {
                cdiv += cdu[p][p];
                ndiv += ndu[p][p];
#pragma unroll
                for (int b = p + 1; b < D; b++) {
                    ccurl += (cdu[p][b] - cdu[b][p]) * (cdu[p][b] - cdu[b][p]);
                    ncurl += (ndu[p][b] - ndu[b][p]) * (ndu[p][b] - ndu[b][p]);
                }
                const double cu = qP[p + 1] / qP[0];
                cvm += cu * cu;
                const double nu = qN[p + 1] / qN[0];
                nvm += nu * nu;
                double gc[D], gn[D];
#pragma unroll
                for (int k = 0; k < D; k++) {
                    gc[k] = gr[0][R::kUG + p * D + k];
                    gn[k] = gr[1][R::kUG + p * D + k];
                }
                const double uPos = interp_minmod<D>(cu, nu, gc, d, w, 1.0);
                const double uNeg = interp_minmod<D>(cu, nu, gn, d, w, -1.0);
                unP += uPos * S[p];
                unN += uNeg * S[p];
                nrm += S[p] * S[p];
            }
            const double cth0 = -(cdiv / sqrt(cdiv * cdiv + ccurl + 4e-2));
            const double nth0 = -(ndiv / sqrt(ndiv * ndiv + ncurl + 4e-2));
            const double cth = (cth0 < 0.) ? 0. : cth0;
            const double nth = (nth0 < 0.) ? 0. : nth0;
            const double tha = 0.5 * (cth + nth);
            double r = qP[0], E = qP[D + 1] / r;
            const double cp_ = r * a.gm1 * (E - 0.5 * cvm);
            r = qN[0];
            E = qN[D + 1] / r;
            const double np_ = r * a.gm1 * (E - 0.5 * nvm);
            double gc[D], gn[D], hc[D], hn[D];
#pragma unroll
            for (int k = 0; k < D; k++) {
                gc[k] = gr[0][R::kRG + k];
                gn[k] = gr[1][R::kRG + k];
                hc[k] = gr[0][R::kPG + k];
                hn[k] = gr[1][R::kPG + k];
            }
            const double rhoP = interp_minmod<D>(qP[0], qN[0], gc, d, w, 1.0);
            const double rhoN = interp_minmod<D>(qP[0], qN[0], gn, d, w, -1.0);
            const double pP = interp_minmod<D>(cp_, np_, hc, d, w, 1.0);
            const double pN = interp_minmod<D>(cp_, np_, hn, d, w, -1.0);
            const double Msq = (unP * unP + unN * unN) / (2. * cavg * cavg * nrm);
            const double MP = unP / (sqrt(nrm) * cavg);
            const double MN = unN / (sqrt(nrm) * cavg);
            const double Minf = 0.2;
            const double mx = (Msq < Minf * Minf) ? Minf * Minf : Msq;
            const double M0 = sqrt((mx < 1.0) ? mx : 1.0);
            const double fa = M0 * (2.0 - M0);
            const double alpha = 3.0 * (5.0 * fa * fa - 4.0) / 16.0;
            const double phalf = pN * (p5pos(MN, alpha) - p5neg(MN, alpha)) - pP * (p5pos(MP, alpha) - p5neg(MP, alpha));
            const double pu = -0.75 * p5pos(MP, alpha) * p5neg(MN, alpha) * (rhoP + rhoN) * cavg * fa * (unN - unP);
            pavg += tha * (pu - 0.5 * phalf);
        }
        double phi = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) phi += (rUa[k] * rai) * S[k];
        rhs[0] = -rhoavg * phi;
#pragma unroll
        for (int k = 0; k < D; k++) rhs[k + 1] = -(rUa[k] * phi + pavg * S[k]);
        rhs[D + 1] = -(rhoavg * Havg * phi);
    }

    // viscous terms (cfd_v0.cpp:2682-2790)
    if (VISC) {
    const double mu = a.mu;
    double dmag = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) dmag += d[k] * d[k];
    dmag = sqrt(dmag);
    const double dmi = 1.0 / dmag;
    double dn[D], du[D][D], dT[D], tau[D][D], tmc[D][D];
    double duP[D][D], duN[D][D], sUP[D], sUN[D];
#pragma unroll
    for (int k = 0; k < D; k++) dn[k] = d[k] * dmi;
    if (ghost) {
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) du[i][j] = (aU[i] - cU[i]) * dn[j] * dmi;
            dT[i] = (aT - cT) * dn[i] * dmi;
        }
    } else {
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                duP[i][j] = gr[0][i * D + j];
                duN[i][j] = gr[1][i * D + j];
                du[i][j] = INTERP_LINEAR(w, duP[i][j], duN[i][j]);
            }
            dT[i] = INTERP_LINEAR(w, gr[0][R::kDT + i], gr[1][R::kDT + i]);
        }
    }
#pragma unroll
    for (int p = 0; p < D; p++) {
        tau[p][p] = 2.0 * du[p][p];
#pragma unroll
        for (int b1 = p + 1; b1 < D + p; b1++) {
            const int b = b1 % D;
            tau[p][b] = mu * (du[p][b] + du[b][p]);
            tau[p][p] -= du[b][b];
        }
        tau[p][p] *= 2.0 / 3.0 * mu;
    }
    const double kcond = a.cp * mu * a.pr_inv;
    if (ghost) {
        double diag = 0;
#pragma unroll
        for (int p = 0; p < D; p++) diag -= du[p][p];
        diag *= mu * 2.0 / 3.0;
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) tmc[i][j] = mu * du[j][i];
            tmc[i][i] += diag;
        }
    } else if constexpr (STORED) {
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++)
                tmc[i][j] = INTERP_LINEAR(w, a.tauMC[(i * D + j) * nt + P], a.tauMC[(i * D + j) * nt + N]);
    } else {
        // both cells' tensors as calc_VIS leaves them, from their dudx
        double tP[D][D], tN[D][D];
        vis_tensors<D>(a, duP, qP, a.smag ? a.smag + P : nullptr, tP, sUP);
        vis_tensors<D>(a, duN, qN, a.smag ? a.smag + N : nullptr, tN, sUN);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) tmc[i][j] = INTERP_LINEAR(w, tP[i][j], tN[i][j]);
    }
    const double Sd = dotD<D>(S, d);
    double delta[D], K[D], dm = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) {
        delta[i] = d[i] * Sm * Sm / Sd;
        dm += delta[i] * delta[i];
        K[i] = S[i] - delta[i];
    }
    dm = sqrt(dm);
    double sU[D];
    if (ghost) {
        double Uf[D];
#pragma unroll
        for (int i = 0; i < D; i++) Uf[i] = INTERP_LINEAR(w, cU[i], aU[i]);
#pragma unroll
        for (int i = 0; i < D; i++) sU[i] = dotD<D>(Uf, tau[i]);
    } else if constexpr (STORED) {
#pragma unroll
        for (int i = 0; i < D; i++) sU[i] = INTERP_LINEAR(w, a.sigmaU[i * nt + P], a.sigmaU[i * nt + N]);
    } else {
#pragma unroll
        for (int i = 0; i < D; i++) sU[i] = INTERP_LINEAR(w, sUP[i], sUN[i]);
    }
#pragma unroll
    for (int i = 0; i < D; i++) {
        const double divT = dotD<D>(tmc[i], S);
        const double lapU = mu * (dm * (aU[i] - cU[i]) * dmi + dotD<D>(K, du[i]));
        rhs[i + 1] += divT + lapU;
    }
    const double lapT = kcond * (dm * (aT - cT) * dmi + dotD<D>(K, dT));
    rhs[D + 1] += dotD<D>(sU, S) + lapT;
    }
    st64x2_once<D, 3>(a.frhs + 4 * (size_t)f, rhs[0], rhs[1]);
    st64x2_once<D, 3>(a.frhs + 4 * (size_t)f + 2, rhs[2], rhs[3]);
    if (D == 3) st64_once<D, 3>(a.frhs + (size_t)4 * a.nf + f, rhs[D + 1]);
}

// tauMC and sigmaU of every cell from the stored dudx (field read-back of a
// solver that does not store them): [D*D][nc] and [D][nc]
template <int D>
__global__ void __launch_bounds__(256) vis_tensor_field(KArgs a, double *tmc_out, double *sU_out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    double du[D][D], q[D + 2], tmc[D][D], sU[D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) du[i][j] = a.dudx[(size_t)c * a.gstride + i * D + j];
    load_q<D>(a, c, q);
    vis_tensors<D>(a, du, q, a.smag ? a.smag + c : nullptr, tmc, sU);
#pragma unroll
    for (int i = 0; i < D; i++) {
        sU_out[(size_t)i * a.nc + c] = sU[i];
#pragma unroll
        for (int j = 0; j < D; j++) tmc_out[(size_t)(i * D + j) * a.nc + c] = tmc[i][j];
    }
}

// ---------------------------------------------------------------------------
// compute_dt / compute_cfl (cfd_v0.cpp:2887-2942) of one cell from its state
// q: the sum over its faces in slot order, then dt_c (MODE 0) or cfl_c
// (MODE 1).  Returns the block's starting value combined with it, so a NaN
// never wins (as `dt_min > dt`).
// ---------------------------------------------------------------------------
template <int D, int MODE>
__device__ __forceinline__ double cfl_cell(const KArgs &a, const int c, const double (&q)[D + 2], double param) {
    double sum = 0.0;
    auto add = [&](const double (&S)[D]) {
        if (MODE == 0) {
            double sM = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) sM += S[k] * S[k];
            sM = sqrt(sM);
            double fl = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) fl += q[k + 1] * S[k] / sM;
            sum += fabs(fl);
        } else {
            double fl = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) fl += q[k + 1] * S[k];
            sum += fabs(fl);
        }
    };
    // slot-ordered sum; the first slots' area vectors fetched together
    // (slot lists carry no ghost flag)
    const int e0 = a.s_off[c], ne = a.s_off[c + 1] - e0;
    constexpr int U = GatherUnroll<D>::value;
    double Sj[U][D];
#pragma unroll
    for (int j = 0; j < U; j++) {
        const int ent = a.s_ent[e0 + (j < ne ? j : ne - 1)];
        const int f = ent >= 0 ? ent : ~ent;
        double sf[D];
        load_S64<D>(a.S, f, sf);
#pragma unroll
        for (int k = 0; k < D; k++) Sj[j][k] = ent >= 0 ? sf[k] : -sf[k];
    }
#pragma unroll
    for (int j = 0; j < U; j++)
        if (j < ne) add(Sj[j]);
    for (int e = e0 + U; e < e0 + ne; e++) {
        const int ent = a.s_ent[e];
        const int f = ent >= 0 ? ent : ~ent;
        double S[D];
        double sf[D];
        load_S64<D>(a.S, f, sf);
#pragma unroll
        for (int k = 0; k < D; k++) S[k] = ent >= 0 ? sf[k] : -sf[k];
        add(S);
    }
    double v = MODE == 0 ? 100.0 : 0.0;
    if (MODE == 0) {
        const double vol = 1.0 / a.vinv[c];
        const double dtc = param * q[0] * 2.0 * vol / sum;
        if (v > dtc) v = dtc;
    } else {
        const double cfl = sum * ((1.0 / q[0]) * a.vinv[c] * param);
        if (v < cfl) v = cfl;
    }
    return v;
}

// Block-wide min (MODE 0) / max (MODE 1) of v into *dst; every thread of the
// block calls it.  The values are never NaN, so min / max are exact in any
// order.
template <int MODE>
__device__ __forceinline__ void block_extremum(double v, double *dst) {
    __shared__ double ext[256 / 64];
    auto pick = [](double x, double o) {
        if (MODE == 0) return x > o ? o : x;
        return x < o ? o : x;
    };
    for (int off = 32; off > 0; off >>= 1) v = pick(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) ext[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = ext[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) r = pick(r, ext[i]);
        *dst = r;
    }
}

// ---------------------------------------------------------------------------
// Residual gather, sponge, RK update (cfd_v0.cpp:2792-2822).  dq starts from
// A_rk * dq (prepare_for_RKstep); zero_dq replaces it by 0 for the first
// stage of a step (prepare_for_timestep zeroed it).  CFL >= 0 (the last
// stage of a step): the step's compute_dt (0) / compute_cfl (1) of the
// updated cells, per-block extrema into a.partial.
// ---------------------------------------------------------------------------
template <int D, bool RES, int CFL>
__global__ void __launch_bounds__(256) cell_update(KArgs a, double A, double B, int zero_dq, double cfl_param) {
    static_assert(!(RES && CFL >= 0), "the residual and the CFL share a.partial");
    const int c = LFM_TID(blockIdx.x);
    __shared__ double red[256 / 64][D + 2];
    double cv = CFL == 0 ? 100.0 : 0.0;
    double rsq[D + 2];
#pragma unroll
    for (int k = 0; k < D + 2; k++) rsq[k] = 0.0;
    if (c < a.nc) {
        const double dt = *a.dt;
        const double vinv = ld64_once<D, 3>(a.vinv + c);
        double dq[D + 2], res[D + 2];
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            dq[k] = zero_dq ? 0.0 : ld64_once<D, 2>(a.dq + (size_t)k * a.nc + c) * A;
            res[k] = 0.0;
        }
        const int e0 = a.g_off[c], ne = a.g_off[c + 1] - e0;
        int gface[kMaxFaces], ngf = 0;
        // owner side adds, neighbour side subtracts (cfd_v0.cpp:2800-2806)
        auto add = [&](const int ent, const double (&fr)[D + 2]) {
            if (ent >= 0) {
                if ((ent & kGhostFaceFlag) && ngf < kMaxFaces) gface[ngf++] = ent & kFaceMask;
#pragma unroll
                for (int k = 0; k < D + 2; k++) {
                    if (RES) res[k] += fr[k];
                    dq[k] += dt * fr[k] * vinv;
                }
            } else {
#pragma unroll
                for (int k = 0; k < D + 2; k++) {
                    if (RES) res[k] -= fr[k];
                    dq[k] -= dt * fr[k] * vinv;
                }
            }
        };
        // the first faces' fluxes fetched together, then added in gather order
        constexpr int U = GatherUnroll<D>::value;
        int ent[U];
        double fr[U][D + 2];
#pragma unroll
        for (int j = 0; j < U; j++) {
            ent[j] = a.g_ent[e0 + (j < ne ? j : ne - 1)];
            const int f = ent[j] >= 0 ? (ent[j] & kFaceMask) : ~ent[j];
            load_frhs<D>(a, f, fr[j]);
        }
#pragma unroll
        for (int j = 0; j < U; j++)
            if (j < ne) add(ent[j], fr[j]);
        for (int e = e0 + U; e < e0 + ne; e++) {
            const int en = a.g_ent[e];
            const int f = en >= 0 ? (en & kFaceMask) : ~en;
            double f1[D + 2];
            load_frhs<D>(a, f, f1);
            add(en, f1);
        }
        double q[D + 2];
        load_q<D>(a, c, q);
        const double sg = ld64_once<D, 3>(a.sigma + c);
        dq[0] += dt * sg * (a.rho_inf - q[0]);
#pragma unroll
        for (int k = 0; k < D; k++) dq[k + 1] += dt * sg * (a.rho_inf * a.U_inf[k] - q[k + 1]);
        dq[D + 1] += dt * sg * (a.rho_inf * a.E_inf - q[D + 1]);
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            q[k] += B * dq[k];
            st64_once<D, 2>(a.dq + (size_t)k * a.nc + c, dq[k]);
        }
        store_q<D>(a.q, a.ntot, c, q);
        // refresh this cell's boundary ghosts for the next stage
        for (int i = 0; i < ngf; i++) {
            const int g = a.fN[gface[i]] - a.nc;
            double qg[D + 2];
            if (!ghost_state<D>(a, a.ghost_role[g], q, qg)) continue;
            store_q<D>(a.q, a.ntot, a.nc + g, qg);
        }
        if constexpr (CFL >= 0) cv = cfl_cell<D, CFL>(a, c, q, cfl_param);
        if (RES) {
#pragma unroll
            for (int k = 0; k < D + 2; k++) {
                a.res_cell[(size_t)k * a.nc + c] = res[k];
                rsq[k] = res[k] * res[k];
            }
        }
    }
    if (RES) {
        // per-block partial sums of squared residuals (fixed order)
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            double v = rsq[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) red[wid][k] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < D + 2; k++) {
                double s = 0.0;
                for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i][k];
                a.partial[(size_t)LFM_BLK() * (D + 2) + k] = s;
            }
        }
    }
    if constexpr (CFL >= 0) block_extremum<CFL>(cv, a.partial + LFM_BLK());
}

// Sum of per-block partials in fixed order, one block.
static __global__ void __launch_bounds__(256) reduce_sum(const double *partial, int nblk, int ncomp, double *out, int accumulate) {
    __shared__ double sh[256];
    for (int k = 0; k < ncomp; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += partial[(size_t)i * ncomp + k];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int st = blockDim.x / 2; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = accumulate ? out[k] + sh[0] : sh[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// compute_dt / compute_cfl (cfd_v0.cpp:2887-2942) as their own launch: per
// cell over its faces in slot order, then min / max.  MODE 0: dt (min, start
// 100), 1: cfl (max, start 0).
// ---------------------------------------------------------------------------
template <int D, int MODE>
__global__ void __launch_bounds__(256) cfl_dt_cells(KArgs a, double param) {
    const int c = LFM_TID(blockIdx.x);
    double v = MODE == 0 ? 100.0 : 0.0;
    if (c < a.nc) {
        double q[D + 2];
        load_q<D>(a, c, q);
        v = cfl_cell<D, MODE>(a, c, q, param);
    }
    block_extremum<MODE>(v, a.partial + LFM_BLK());
}

// Final min / max over the per-block values by one block of NT threads,
// each with 16 loads in flight per round, so the latency of a few rounds
// (65 536 partials on 16.8 M cells: 4 with 1024 threads) bounds it rather
// than a long dependent chain.  Min / max are exact in any order.
template <int MODE, int NT>
__device__ __forceinline__ void cfl_dt_final_block(const double *partial, int nblk, double *out, double *dt_dst) {
    auto pick = [](double x, double o) {
        if (MODE == 0) return x > o ? o : x;
        return x < o ? o : x;
    };
    constexpr int kLoads = 16;
    double v = MODE == 0 ? 100.0 : 0.0;
    for (int i0 = 0; i0 < nblk; i0 += kLoads * NT) {
        double o[kLoads];
#pragma unroll
        for (int j = 0; j < kLoads; j++) {
            const int i = i0 + j * NT + (int)threadIdx.x;
            o[j] = i < nblk ? partial[i] : v;
        }
#pragma unroll
        for (int j = 0; j < kLoads; j++) v = pick(v, o[j]);
    }
    for (int off = 32; off > 0; off >>= 1) v = pick(v, __shfl_down(v, off, 64));
    __shared__ double w[NT / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = w[0];
        for (int i = 1; i < NT / 64; i++) r = pick(r, w[i]);
        *out = r;
        if (dt_dst) *dt_dst = r;
    }
}

// the final reduction as a launch of its own: one block of kFinalBlock
constexpr int kFinalBlock = 1024;
template <int MODE>
static __global__ void __launch_bounds__(kFinalBlock) cfl_dt_final(const double *partial, int nblk, double *out,
                                                                  double *dt_dst) {
    cfl_dt_final_block<MODE, kFinalBlock>(partial, nblk, out, dt_dst);
}

// conservation integrals sum_c V_c q_c, per-block partials (FP64)
template <int D>
__global__ void __launch_bounds__(256) integrals64(KArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double red[256 / 64][D + 2];
    double v[D + 2];
#pragma unroll
    for (int k = 0; k < D + 2; k++) v[k] = c < a.nc ? a.q[q64_index(k, c, a.ntot)] / a.vinv[c] : 0.0;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < D + 2; k++) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[wid][k] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            double s = 0.0;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i][k];
            a.partial[(size_t)blockIdx.x * (D + 2) + k] = s;
        }
    }
}

static __global__ void __launch_bounds__(256) fill_zero(double *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

template <int D>
__global__ void __launch_bounds__(256) count_nonfinite(const double *q, int nc, int ntot, unsigned long long *cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int bad = 0;
#pragma unroll
    for (int k = 0; k < D + 2; k++) bad += !isfinite(q[q64_index(k, c, ntot)]);
    if (bad) atomicAdd(cnt, (unsigned long long)bad);
}

}  // namespace lfm
