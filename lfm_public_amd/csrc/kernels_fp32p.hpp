// kernels_fp32p.hpp -- FP32 time step on deviations from the freestream.
//
// State: q' = q - q_inf in FP32 (q_inf = rho_inf, rho_inf U_inf, rho_inf E_inf
// held as constants).  Every face flux is evaluated as F(q_f).S - F(q_inf).S
// with the deviation formed algebraically, never by subtracting two
// freestream-sized numbers:
//   u' = (m' - u_inf rho') / rho
//   p' = (g-1) [ (rhoE)' - 1/2 (rho' |u|^2 + rho_inf u'.(2 u_inf + u')) ]
//   Rpsi' = (p' - Rpsi_inf rho') / rho
// Subtracting F(q_inf).S face by face leaves every closed cell's sum unchanged
// in exact arithmetic (sum of S over a closed cell is zero) and removes the
// FP32 closure error of the S vectors times the freestream flux.  Green-Gauss
// gradients are taken of the deviations for the same reason.  dt, its
// min-reduction, the CFL and the conservation integrals stay FP64.
//
// M1 (cfd_v0.cpp:2589-2675): both reconstructed states are the same linear
// interpolation (the reference forms rho- with weight 1-w on the swapped pair),
// so the central-upwind combination [psi+ F+ - psi- F- + psi+ psi- (q- - q+)]
// / (psi+ - psi-) equals F+ exactly; this path evaluates F+ directly and so
// carries none of the rounding-level dissipation the FP64 form contains.
#pragma once

#include "solver_state.hpp"

namespace lfm {

struct KArgs32 {
    int nc, nsm0, ng, nf, ntot;
    int ng_bnd;  // boundary ghosts [nc, nc+ng_bnd); halo ghosts after them
    const int32_t *fP, *fN;
    const float *S;  // [nf][D]: a face's area vector as one record
    const float *d, *w;
    const int32_t *g_off, *g_ent;
    const int32_t *ghost_cell, *ghost_role;
    const float *vinv;
    const float *sigma;  // nullptr where every cell's sponge sigma is zero
    const float *smag;  // [nc] Smagorinsky constants, nullptr when laminar
    float *q, *dq, *res_cell;
    // per-cell gradient records (GradRec32): dudx points at the records,
    // dTdx / rho_grad / p_grad into them (component stride 1, cell stride
    // gstride); tauMC / sigmaU (partitions only) and a U_grad that is not
    // dudx are [component][ntot]
    float *dudx, *dTdx, *tauMC, *sigmaU;
    float *rho_grad, *p_grad, *U_grad;
    int gstride;
    // face fluxes: components 0-3 as one 16-byte record per face ([nf][4]),
    // then (3D) component 4 as [nf]: one vector load per gathered face
    float *frhs;
    const double *dt;
    double *partial;
    const double *vol_d;  // FP64 cell volumes (conservation integrals)
    double *d_cfl, *d_dtmin;  // this solver's FP64 CFL / dt-min scalars
    int nblk_cells;           // blocks of 256 over the cells (reduction partials)
    // per-cell face slots of the one-launch stage and the 3D gradient kernel
    // (the first GatherUnroll gather entries of each cell, stored as
    // slot_at places them): the face's other cell (~other where the cell is
    // the face's N; the cell itself past its last entry), its area vector
    // (face orientation) and weight; nullptr until built
    const int32_t *sl_nb;
    const float *sl_S, *sl_w;
    // the second state buffer, in a batch's copy of a member's arguments
    // (the solver's own copy leaves it null)
    float *q_alt;
    float gamma, gm1, gog, R_inv, mu, cp, pr_inv;
    float rho_inf, p_inf, T_inf, Rpsi_inf, E_inf, H_inf, G_inf, c_inf;
    float uinf[3], minf[3];
    double rho_inf_d, m_inf_d[3], rhoE_inf_d;
};

#define LERP32(w, a, b) ((w) * (a) + (1.0f - (w)) * (b))

// FP32 reciprocal, quotient, square root and reciprocal square root on the
// hardware instructions (v_rcp_f32, v_sqrt_f32, v_rsq_f32: 1 ulp).  Spelled
// as target builtins so that every kernel instantiating a body, single or
// batched, selects the same instructions: the lowering of a plain `/` or
// sqrtf depends on whether its operands are wave-uniform, which differs
// between the two launch forms.
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float div32(float n, float d) { return n * __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ float sqrt32(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq32(float x) { return __builtin_amdgcn_rsqf(x); }

// face f's area vector (one 8-byte load in 2D)
template <int D>
__device__ __forceinline__ void load_S32(const float *S, int f, float (&s)[D]) {
    if constexpr (D == 2) {
        const float2 v = reinterpret_cast<const float2 *>(S)[f];
        s[0] = v.x;
        s[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < D; k++) s[k] = S[(size_t)f * D + k];
    }
}

template <int D>
struct Prim32 {
    float rp, rho, ri, up[D], u[D], pp, Rp, R;
};

template <int D>
__device__ __forceinline__ Prim32<D> prim32(const KArgs32 &a, const float *q) {
    Prim32<D> s;
    s.rp = q[0];
    s.rho = a.rho_inf + s.rp;
    s.ri = rcp32(s.rho);
    float u2 = 0.f, du2 = 0.f;
#pragma unroll
    for (int k = 0; k < D; k++) {
        s.up[k] = (q[k + 1] - a.uinf[k] * s.rp) * s.ri;
        s.u[k] = a.uinf[k] + s.up[k];
        u2 += s.u[k] * s.u[k];
        du2 += s.up[k] * (2.f * a.uinf[k] + s.up[k]);
    }
    s.pp = a.gm1 * (q[D + 1] - 0.5f * (s.rp * u2 + a.rho_inf * du2));
    s.Rp = (s.pp - a.Rpsi_inf * s.rp) * s.ri;
    s.R = a.Rpsi_inf + s.Rp;
    return s;
}

// FP32 state layout: components 0-3 of a cell (or ghost) i as one 16-byte
// record, q[4 i + k], then component 4 (the 3D energy) as q[4 ntot + i]: a
// gathered neighbour state is one vector load (two in 3D).
__host__ __device__ __forceinline__ size_t q32_index(int k, size_t i, size_t ntot) {
    return k < 4 ? 4 * i + (size_t)k : 4 * ntot + i;
}

// loads and stores of data a kernel touches once, marked non-temporal so
// that they do not displace the gathered neighbour states in L2: level 1 the
// face slots, 2 the RK register, 3 the other streams (new states, face
// lists and geometry of the face kernel, face fluxes, gradient records,
// 1/V and sponge).  3D uses every level (256^3 box 3.70 -> 3.35 ms per
// step, 4.2 M-cell O-grid 2.75 -> 2.51).  2D uses level 1 only: a 2D
// solve's arrays fit in the 256 MB Infinity Cache, where the next kernel
// finds what this one wrote unless it was written non-temporally (1 M-cell
// annulus 0.32 -> 0.40 ms per step with every level).
template <int D>
constexpr int kStreamNT = D == 3 ? 3 : 1;
using f4v = float __attribute__((ext_vector_type(4)));
template <int D, int LEVEL, class T>
__device__ __forceinline__ T ld_once(const T *p) {
    if constexpr (kStreamNT<D> >= LEVEL) return __builtin_nontemporal_load(p);
    return *p;
}
template <int D, int LEVEL, class T>
__device__ __forceinline__ void st_once(T *p, T v) {
    if constexpr (kStreamNT<D> >= LEVEL)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
// a 16-byte record
template <int D, int LEVEL>
__device__ __forceinline__ void st4_once(float *p, float x, float y, float z, float w) {
    const f4v v = {x, y, z, w};
    st_once<D, LEVEL>(reinterpret_cast<f4v *>(p), v);
}

template <int D>
__device__ __forceinline__ void load_q32(const KArgs32 &a, int i, float *q) {
    const float4 v = reinterpret_cast<const float4 *>(a.q)[i];
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
    q[3] = v.w;
    if (D == 3) q[4] = a.q[(size_t)4 * a.ntot + i];
}

template <int D>
__device__ __forceinline__ void store_q32(float *qv, size_t ntot, int i, const float *q) {
    st4_once<D, 3>(qv + 4 * (size_t)i, q[0], q[1], q[2], q[3]);
    if (D == 3) st_once<D, 3>(qv + 4 * ntot + i, q[4]);
}

template <int D>
__device__ __forceinline__ float dot32(const float *x, const float *y) {
    float s = x[0] * y[0];
#pragma unroll
    for (int k = 1; k < D; k++) s += x[k] * y[k];
    return s;
}

__device__ __forceinline__ float sgn32(float x) { return x < 0.0f ? -1.0f : 1.0f; }

// interp_minmod (api/cfdv0_solver.h:311-370) on deviations: phi' differences
// equal phi differences, and the weights sum to one, so phi_inf is added back
// by the caller.
template <int D>
__device__ __forceinline__ float minmod32(float cphi, float aphi, const float *grad, const float *d, float wl,
                                          float flux) {
    const float gradf = aphi - cphi + 1.0e-30f;
    float gradcf = 0.0f;
#pragma unroll
    for (int k = 0; k < D; k++) gradcf += d[k] * grad[k];
    float r;
    if (fabsf(gradcf) >= 1000.0f * fabsf(gradf))
        r = 2.0f * 1000.0f * sgn32(gradcf) * sgn32(gradf) - 1.0f;
    else
        r = 2.0f * div32(gradcf, gradf) - 1.0f;
    const float m1 = (1.0f < r) ? 1.0f : r;
    const float lim = (m1 < 0.0f) ? 0.0f : m1;
    const float wt = lim * wl + (1.0f - lim) * (1.0f + flux) * 0.5f;
    return wt * cphi + (1.0f - wt) * aphi;
}

__device__ __forceinline__ float p5pos32(float M, float alpha) {
    const float M2Pos = 0.25f * (M + 1.0f) * (M + 1.0f);
    const float M2Neg = -0.25f * (M - 1.0f) * (M - 1.0f);
    if (fabsf(M) < 1.f) return M2Pos * ((2.0f - M) - 16.0f * alpha * M * M2Neg);
    return div32(0.5f * (M + fabsf(M)), M);
}

__device__ __forceinline__ float p5neg32(float M, float alpha) {
    const float M2Pos = 0.25f * (M + 1.0f) * (M + 1.0f);
    const float M2Neg = -0.25f * (M - 1.0f) * (M - 1.0f);
    if (fabsf(M) < 1.f) return M2Neg * ((-2.0f - M) + 16.0f * alpha * M * M2Pos);
    return div32(0.5f * (M - fabsf(M)), M);
}

// ---------------------------------------------------------------------------
// ghost states (cfd_v0.cpp:1010-1133) in deviation form
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ bool ghost_state32(const KArgs32 &a, int role, const float *qi, float *qg) {
    if (role == 1) {  // wall: rho, -rhoU, rhoE
        qg[0] = qi[0];
#pragma unroll
        for (int k = 0; k < D; k++) qg[k + 1] = -qi[k + 1] - 2.f * a.minf[k];
        qg[D + 1] = qi[D + 1];
        return true;
    }
    if (role != 2 && role != 3) return false;  // unset patch: q = 0, i.e. q' = -q_inf
    const Prim32<D> s = prim32<D>(a, qi);
    if (role == 2) {  // inlet: T = T_inf, U = U_inf, p = p_int
        const float rpg = s.rp + div32(s.rho * (s.Rp * a.R_inv), a.T_inf);
        qg[0] = rpg;
        float u2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            qg[k + 1] = rpg * a.uinf[k];
            u2 += a.uinf[k] * a.uinf[k];
        }
        qg[D + 1] = div32(s.pp, a.gm1) + 0.5f * rpg * u2;
    } else {  // outlet: p = p_inf, U = U_int, rho = rho_int p_inf / p_int
        const float pint = a.p_inf + s.pp;
        const float rpg = div32(s.rp * a.p_inf - a.rho_inf * s.pp, pint);
        const float rhog = a.rho_inf + rpg;
        float u2 = 0.f, du2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            qg[k + 1] = rpg * a.uinf[k] + rhog * s.up[k];
            u2 += s.u[k] * s.u[k];
            du2 += s.up[k] * (2.f * a.uinf[k] + s.up[k]);
        }
        qg[0] = rpg;
        qg[D + 1] = 0.5f * (rpg * u2 + a.rho_inf * du2);
    }
    return true;
}

// ghost states (cfd_v0.cpp:1010-1133) in deviation form; the fused update
// refreshes ghosts, so this runs only after the state is set from outside
template <int D>
__device__ __forceinline__ void bc_fill32_body(const KArgs32 &a, const int g) {
    if (g >= a.ng) return;
    float qi[D + 2], qg[D + 2];
    load_q32<D>(a, a.ghost_cell[g], qi);
    if (!ghost_state32<D>(a, a.ghost_role[g], qi, qg)) return;
    const int gi = a.nc + g;
    store_q32<D>(a.q, a.ntot, gi, qg);
}

// ---------------------------------------------------------------------------
// Green-Gauss of deviations + viscous tensors, per cell
// ---------------------------------------------------------------------------
// tauMC and sigmaU of one cell from its completed dudx and velocity (the
// calc_VIS refresh, Smagorinsky stress when smag is given).  Used where they
// are stored and where a face recomputes them.
template <int D>
__device__ __forceinline__ void vis_tensors32(const KArgs32 &a, const float (&dudx)[D][D], const float (&u)[D],
                                              const float *smag_p, float (&tmc)[D][D], float (&sU)[D]) {
    const float mu = a.mu;
    float tau[D][D];
#pragma unroll
    for (int p = 0; p < D; p++) {
        tau[p][p] = 2.0f * dudx[p][p];
#pragma unroll
        for (int b1 = p + 1; b1 < D + p; b1++) {
            const int b = b1 % D;
            tau[p][b] = mu * (dudx[p][b] + dudx[b][p]);
            tau[p][p] -= dudx[b][b];
        }
        tau[p][p] *= 2.0f / 3.0f * mu;
    }
    float diag = 0.f;
#pragma unroll
    for (int p = 0; p < D; p++) diag -= dudx[p][p];
    diag *= mu * (2.0f / 3.0f);
#pragma unroll
    for (int i = 0; i < D; i++) {
#pragma unroll
        for (int j = 0; j < D; j++) tmc[i][j] = mu * dudx[j][i];
        tmc[i][i] += diag;
    }
    if (smag_p) {  // Smagorinsky stress (cfd_v0.cpp:1668-1684)
        const float smag = *smag_p;
        float strain = 0.f, divu = 0.f;
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                const float sij = 0.5f * (dudx[i][j] + dudx[j][i]);
                strain += sij * sij;
            }
            divu += dudx[i][i];
        }
        strain = sqrt32(2.f * strain);
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) tmc[i][j] -= smag * strain * (0.5f * (dudx[i][j] + dudx[j][i]));
            tmc[i][i] += smag * strain * divu * (1.f / 3.f);
        }
    }
#pragma unroll
    for (int i = 0; i < D; i++) sU[i] = dot32<D>(u, tau[i]);
}

// One cell's gradients as one record, padded to whole 8-byte words: dudx
// (D*D), dTdx (D), then with RG (the M2AUSM scheme) the rho and p gradients
// (D each); ghosts' records stay zero.  A gathered neighbour's gradients are
// a few 8-byte vector loads.
template <int D, bool RG>
struct GradRec32 {
    static constexpr int kDT = D * D;
    static constexpr int kRG = kDT + D;
    static constexpr int kPG = kRG + D;
    static constexpr int kN = RG ? kPG + D : kRG;
    static constexpr int kStride = (kN + 1) / 2 * 2;
};
inline int grad_record_stride(int D, bool rg) {
    const int n = D * D + D + (rg ? 2 * D : 0);
    return (n + 1) / 2 * 2;
}

// Gather entries whose loads the FP32 gradient and update kernels issue
// together before any arithmetic (the gradient kernel takes any further
// entries one at a time, the update in further groups of the same size).
// Fewer in flight is slower: O-grid gradients 0.97 -> 1.54 ms per step
// with three, 1.89 with two.
template <int D>
struct GradUnroll32 {
    static constexpr int value = GatherUnroll<D>::value;
};
template <int D>
struct UpdateUnroll32 {
    static constexpr int value = GatherUnroll<D>::value;
};

// Face slot storage: for each group of 64 consecutive cells (one wave's
// cells), slot j's field f (NF fields per slot) at
// [c / 64][j][f][c % 64], so that a wave reads one contiguous block of each
// slot array instead of a separate strided stream per slot and field
// (kSlotBlocked; otherwise [j][f][c]).  Arrays hold whole groups.
constexpr bool kSlotBlocked = true;
template <int D, int NF>
__device__ __forceinline__ size_t slot_at(int nc, int j, int f, int c) {
    if constexpr (kSlotBlocked)
        return ((size_t)(c >> 6) * (GatherUnroll<D>::value * NF) + (size_t)(j * NF + f)) * 64 + (size_t)(c & 63);
    return ((size_t)j * NF + f) * (size_t)nc + c;
}
inline size_t slot_array_len(int nc, int U, int NF) { return (size_t)(nc + 63) / 64 * 64 * U * NF; }

// gather entry j of cell c (entry e of the list): its side, other cell,
// weight and area vector.  SL: from the per-cell face slots for the first
// GatherUnroll entries (streamed, no face index to chase); otherwise, and
// beyond them, from the gather list and the face arrays.
template <int D, bool SL>
__device__ __forceinline__ void gather_entry32(const KArgs32 &a, const int c, const int e, const int j, bool &isP,
                                               int &other, float &w, float (&S)[D]) {
    if (SL && j < GatherUnroll<D>::value) {
        const int key = ld_once<D, 1>(a.sl_nb + slot_at<D, 1>(a.nc, j, 0, c));
        isP = key >= 0;
        other = isP ? key : ~key;
        w = ld_once<D, 1>(a.sl_w + slot_at<D, 1>(a.nc, j, 0, c));
#pragma unroll
        for (int k = 0; k < D; k++) S[k] = ld_once<D, 1>(a.sl_S + slot_at<D, D>(a.nc, j, k, c));
        return;
    }
    const int ent = a.g_ent[e];
    isP = ent >= 0;
    const int f = isP ? (ent & kFaceMask) : ~ent;
    other = isP ? a.fN[f] : a.fP[f];
    w = a.w[f];
    load_S32<D>(a.S, f, S);
}

// Green-Gauss sums of one cell over its faces in gather order; returns the
// cell's own primitive state
template <int D, int GRAD, bool VISC, bool SL>
__device__ __forceinline__ Prim32<D> grad32_sums(const KArgs32 &a, const int c, float (&dudx)[D][D],
                                                  float (&dTdx)[D], float (&rg)[D], float (&pg)[D],
                                                  float (&ug)[D][D]) {
    const float vinv = a.vinv[c];
#pragma unroll
    for (int i = 0; i < D; i++) {
        dTdx[i] = rg[i] = pg[i] = 0.f;
#pragma unroll
        for (int j = 0; j < D; j++) dudx[i][j] = ug[i][j] = 0.f;
    }
    float qc0[D + 2];
    load_q32<D>(a, c, qc0);
    const Prim32<D> own = prim32<D>(a, qc0);
    // one face's contribution (the face values use the face's own P/N order)
    auto add = [&](const bool isP, const float w, const float (&S)[D], const float (&qo)[D + 2]) {
        const Prim32<D> so = prim32<D>(a, qo);
        const Prim32<D> &sp = isP ? own : so;
        const Prim32<D> &sn = isP ? so : own;
        float sv[D];
#pragma unroll
        for (int k = 0; k < D; k++) sv[k] = isP ? S[k] * vinv : -S[k] * vinv;
        float fU[D];
#pragma unroll
        for (int k = 0; k < D; k++) fU[k] = LERP32(w, sp.up[k], sn.up[k]);
        const float fT = VISC ? LERP32(w, sp.Rp, sn.Rp) * a.R_inv : 0.f;
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) dudx[i][j] += fU[i] * sv[j];
            if (VISC) dTdx[i] += fT * sv[i];
        }
        if (GRAD == 2) {
            const float rf = LERP32(w, sp.rp, sn.rp);
            const float pf = LERP32(w, sp.pp, sn.pp);
#pragma unroll
            for (int i = 0; i < D; i++) {
                pg[i] += pf * sv[i];
                rg[i] += rf * sv[i];
#pragma unroll
                for (int j = 0; j < D; j++) ug[j][i] += fU[j] * sv[i];
            }
        }
    };
    // every entry in the slots (3D): the gather list's offsets are not read,
    // a placeholder is known by its other cell, the cell itself
    constexpr bool kAll = SL && GatherUnroll<D>::value >= kMaxFaces;
    const int e0 = kAll ? 0 : a.g_off[c], ne = kAll ? kMaxFaces : a.g_off[c + 1] - e0;
    // the first faces: every load issued before the sums (a short list
    // repeats its last face, whose terms are then not added)
    constexpr int U = GradUnroll32<D>::value;
    bool isP[U], valid[U];
    float w[U], S[U][D], qo[U][D + 2];
#pragma unroll
    for (int j = 0; j < U; j++) {
        // past the cell's last entry: the last entry again (list) or the
        // slot's placeholder (slots), not added
        const int jj = j < ne ? j : ne - 1;
        int other;
        gather_entry32<D, SL>(a, c, e0 + jj, SL ? j : jj, isP[j], other, w[j], S[j]);
        valid[j] = kAll ? !(isP[j] && other == c) : j < ne;
        load_q32<D>(a, other, qo[j]);
    }
#pragma unroll
    for (int j = 0; j < U; j++)
        if (valid[j]) add(isP[j], w[j], S[j], qo[j]);
    for (int e = e0 + U; e < e0 + ne; e++) {
        bool p;
        int other;
        float wf, Sf[D], q[D + 2];
        gather_entry32<D, SL>(a, c, e, e - e0, p, other, wf, Sf);
        load_q32<D>(a, other, q);
        add(p, wf, Sf, q);
    }
    return own;
}

// every thread of the block must call this (the block's records are written
// through LDS); c = the block's first cell + threadIdx.x
template <int D, int GRAD, bool VISC, bool RG, bool SL = false>
__device__ __forceinline__ void grad_vis32_body(const KArgs32 &a, const int c) {
    using R = GradRec32<D, RG>;
    const bool valid = c < a.nc;
    float r[R::kStride];
#pragma unroll
    for (int k = 0; k < R::kStride; k++) r[k] = 0.f;
    if (valid) {
        float dudx[D][D], dTdx[D], rg[D], pg[D], ug[D][D];
        const Prim32<D> own = grad32_sums<D, GRAD, VISC, SL>(a, c, dudx, dTdx, rg, pg, ug);
        // tauMC / sigmaU stored only where a halo payload needs them (otherwise
        // each face recomputes them from dudx); U_grad equals dudx bit for bit
        // (same face velocity, same sums) and shares its storage unless a halo
        // payload must leave it zero on the ghosts
        const size_t nt = a.ntot;
        if (VISC && a.tauMC) {
            float tmc[D][D], sU[D];
            vis_tensors32<D>(a, dudx, own.u, a.smag ? a.smag + c : nullptr, tmc, sU);
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
            for (int j = 0; j < D; j++) r[i * D + j] = dudx[i][j];
            r[R::kDT + i] = VISC ? dTdx[i] : 0.f;
            if (RG && GRAD == 2) {
                r[R::kRG + i] = rg[i];
                r[R::kPG + i] = pg[i];
            }
        }
        if (GRAD == 2 && a.U_grad != a.dudx)
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) a.U_grad[(i * D + j) * nt + c] = ug[i][j];
        if (!VISC) {
            // dTdx not formed: the formed components only (the record keeps
            // whatever a full calc_VIS left in the dTdx slots)
            float *rec = a.dudx + (size_t)c * R::kStride;
#pragma unroll
            for (int i = 0; i < D; i++) {
#pragma unroll
                for (int j = 0; j < D; j++) rec[i * D + j] = r[i * D + j];
                if (RG && GRAD == 2) {
                    rec[R::kRG + i] = r[R::kRG + i];
                    rec[R::kPG + i] = r[R::kPG + i];
                }
            }
        }
    }
    if (VISC) {
        // the block's records are contiguous from its first cell (a multiple
        // of 256, so 16-byte aligned): staged in LDS, then written as
        // consecutive 16-byte words and a tail of single words (without
        // minmod the rho / p slots are written zero, as they stay)
        constexpr int S = R::kStride;
        __shared__ __align__(16) float st[256 * S];
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < S / 2; k++)
            reinterpret_cast<float2 *>(st)[t * (S / 2) + k] = make_float2(r[2 * k], r[2 * k + 1]);
        __syncthreads();
        const int c0 = c - t;
        const int nrec = a.nc - c0 < (int)blockDim.x ? a.nc - c0 : (int)blockDim.x;
        float *dst = a.dudx + (size_t)c0 * S;
        const int n = nrec * S, n4 = n / 4;
        for (int j = t; j < n4; j += blockDim.x)
            st_once<D, 3>(reinterpret_cast<f4v *>(dst) + j, reinterpret_cast<const f4v *>(st)[j]);
        for (int j = 4 * n4 + t; j < n; j += blockDim.x) dst[j] = st[j];
    }
}

// ---------------------------------------------------------------------------
// Where the face kernel reads the cells' gradients (the solver's arrays,
// indexed by cell)
// ---------------------------------------------------------------------------
template <int D, bool RG>
struct GradGlobal32 {
    const KArgs32 &a;
    using R = GradRec32<D, RG>;
    // the gradient record of cell c
    __device__ __forceinline__ void rec(int c, float (&r)[R::kStride]) const {
        if constexpr (R::kStride % 4 == 0) {
            // records of whole 16-byte words (3D without minmod: 48 B)
            const float4 *p = reinterpret_cast<const float4 *>(a.dudx + (size_t)c * R::kStride);
#pragma unroll
            for (int k = 0; k < R::kStride / 4; k++) {
                const float4 v = p[k];
                r[4 * k] = v.x;
                r[4 * k + 1] = v.y;
                r[4 * k + 2] = v.z;
                r[4 * k + 3] = v.w;
            }
        } else {
            const float2 *p = reinterpret_cast<const float2 *>(a.dudx + (size_t)c * R::kStride);
#pragma unroll
            for (int k = 0; k < R::kStride / 2; k++) {
                const float2 v = p[k];
                r[2 * k] = v.x;
                r[2 * k + 1] = v.y;
            }
        }
    }
    // U_grad: dudx itself, or its own [component][ntot] array
    __device__ __forceinline__ float ug(const float (&r)[R::kStride], int c, int i, int j) const {
        return a.U_grad == a.dudx ? r[i * D + j] : a.U_grad[(size_t)(i * D + j) * a.ntot + c];
    }
};

// ---------------------------------------------------------------------------
// face fluxes minus the freestream flux, one thread per face
// ---------------------------------------------------------------------------
// the flux of face f from its cells' states and primitives (qP, sp of P;
// qN, sn of N), whoever loaded them: the face kernel loads both, the
// one-launch stage holds its own cell's and loads the other
template <int D, int SOLVER, bool VISC, bool STORED, class GA>
__device__ __forceinline__ void face_flux32_states(const KArgs32 &a, const int f, const int P, const int N,
                                                   const float (&qP)[D + 2], const float (&qN)[D + 2],
                                                   const Prim32<D> &sp, const Prim32<D> &sn,
                                                   const float (&S)[D], const float w, const GA &ga,
                                                   const int gP, const int gN, float (&rhs)[D + 2]) {
    using R = GradRec32<D, SOLVER == 2>;
    float gr[2][R::kStride];  // gradient records of P and N
    if (VISC || SOLVER == 2) {
        ga.rec(gP, gr[0]);
        ga.rec(gN, gr[1]);
    }
    // is_ghost: boundary ghosts only; halo ghosts carry the neighbour's tensors
    const bool ghost = (unsigned)(N - a.nc) < (unsigned)a.ng_bnd;
    const size_t nt = a.ntot;
    float d[D];
#pragma unroll
    for (int k = 0; k < D; k++) d[k] = ld_once<D, 3>(a.d + (size_t)k * a.nf + f);
    float Sm2 = 0.f;
#pragma unroll
    for (int k = 0; k < D; k++) Sm2 += S[k] * S[k];
    const float Sm = sqrt32(Sm2);
    float phi_inf = 0.f;
#pragma unroll
    for (int k = 0; k < D; k++) phi_inf += a.uinf[k] * S[k];

    if (SOLVER == 0) {
        // reconstructed state (cfd_v0.cpp:2590-2637): conservatives, Rpsi
        // (hence p and e = Rpsi/(g-1)) interpolated linearly
        const float rp = LERP32(w, qP[0], qN[0]);
        const float rho = a.rho_inf + rp;
        const float ri = rcp32(rho);
        float mp[D], up[D];
        float phip = 0.f, du2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            mp[k] = LERP32(w, qP[k + 1], qN[k + 1]);
            up[k] = (mp[k] - a.uinf[k] * rp) * ri;
            phip += up[k] * S[k];
            du2 += up[k] * (2.f * a.uinf[k] + up[k]);
        }
        const float phi = phi_inf + phip;
        const float Rp = LERP32(w, sp.Rp, sn.Rp);
        const float pp = rp * a.Rpsi_inf + rho * Rp;
        // G = rhoE + p = rho (g/(g-1) Rpsi + |u|^2/2), e = Rpsi/(g-1) per cell
        float uinf2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) uinf2 += a.uinf[k] * a.uinf[k];
        const float Gdev = rp * (a.gog * a.Rpsi_inf + 0.5f * uinf2) + rho * (a.gog * Rp + 0.5f * du2);
        float mS = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) mS += mp[k] * S[k];
        rhs[0] = -mS;
#pragma unroll
        for (int k = 0; k < D; k++) rhs[k + 1] = -(phip * a.minf[k] + phi * mp[k] + pp * S[k]);
        rhs[D + 1] = -(phip * a.G_inf + phi * Gdev);
    } else {
        // midpoint averages (cfd_v0.cpp:1968-2023)
        const float rp = 0.5f * (qP[0] + qN[0]);
        const float rho = a.rho_inf + rp;
        const float ri = rcp32(rho);
        float mp[D], up[D];
        float phip = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            mp[k] = 0.5f * (qP[k + 1] + qN[k + 1]);
            up[k] = (mp[k] - a.uinf[k] * rp) * ri;
            phip += up[k] * S[k];
        }
        const float phi = phi_inf + phip;
        const float Rp = 0.5f * (sp.Rp + sn.Rp);
        float pp = rp * a.Rpsi_inf + rho * Rp;
        // H = rhoE/rho + Rpsi = g/(g-1) Rpsi + |u|^2/2 per cell, averaged
        float duP = 0.f, duN = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) {
            duP += sp.up[k] * (2.f * a.uinf[k] + sp.up[k]);
            duN += sn.up[k] * (2.f * a.uinf[k] + sn.up[k]);
        }
        const float Hp = 0.5f * ((a.gog * sp.Rp + 0.5f * duP) + (a.gog * sn.Rp + 0.5f * duN));
        const float Gdev = rp * a.H_inf + rho * Hp;
        if (SOLVER == 2) {
            const float cP = sqrt32(a.gamma * sp.R), cN = sqrt32(a.gamma * sn.R);
            const float cavg = 0.5f * (cP + cN);
            float cdiv = 0.f, ndiv = 0.f, ccurl = 0.f, ncurl = 0.f, unP = 0.f, unN = 0.f;
            float cdu[D][D], ndu[D][D];
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
                float gc[D], gn[D];
#pragma unroll
                for (int k = 0; k < D; k++) {
                    gc[k] = ga.ug(gr[0], gP, p, k);
                    gn[k] = ga.ug(gr[1], gN, p, k);
                }
                unP += minmod32<D>(sp.up[p], sn.up[p], gc, d, w, 1.0f) * S[p];
                unN += minmod32<D>(sp.up[p], sn.up[p], gn, d, w, -1.0f) * S[p];
            }
            const float cth0 = -cdiv * rsq32(cdiv * cdiv + ccurl + 4e-2f);
            const float nth0 = -ndiv * rsq32(ndiv * ndiv + ncurl + 4e-2f);
            const float tha = 0.5f * (((cth0 < 0.f) ? 0.f : cth0) + ((nth0 < 0.f) ? 0.f : nth0));
            float gc[D], gn[D], hc[D], hn[D];
#pragma unroll
            for (int k = 0; k < D; k++) {
                gc[k] = gr[0][R::kRG + k];
                gn[k] = gr[1][R::kRG + k];
                hc[k] = gr[0][R::kPG + k];
                hn[k] = gr[1][R::kPG + k];
            }
            const float rhoPos = a.rho_inf + minmod32<D>(sp.rp, sn.rp, gc, d, w, 1.0f);
            const float rhoNeg = a.rho_inf + minmod32<D>(sp.rp, sn.rp, gn, d, w, -1.0f);
            const float pPd = minmod32<D>(sp.pp, sn.pp, hc, d, w, 1.0f);
            const float pNd = minmod32<D>(sp.pp, sn.pp, hn, d, w, -1.0f);
            const float unPf = phi_inf + unP, unNf = phi_inf + unN;
            const float rSc = rcp32(Sm * cavg);
            const float MP = unPf * rSc;
            const float MN = unNf * rSc;
            const float Msq = 0.5f * (MP * MP + MN * MN);
            const float Minf = 0.2f;
            const float mx = (Msq < Minf * Minf) ? Minf * Minf : Msq;
            const float M0 = sqrt32((mx < 1.0f) ? mx : 1.0f);
            const float fa = M0 * (2.0f - M0);
            const float alpha = 3.0f * (5.0f * fa * fa - 4.0f) / 16.0f;
            const float dN = p5pos32(MN, alpha) - p5neg32(MN, alpha);
            const float dP = p5pos32(MP, alpha) - p5neg32(MP, alpha);
            // p_half = pN dN - pP dP = p_inf (dN - dP) + pN' dN - pP' dP.  For
            // |M| < 1, P5+(M) - P5-(M) = M (3 - M^2)/2 + 2 alpha M (1 - M^2)^2, so
            // dN - dP is formed as (MN - MP) times its divided difference, with
            // MN - MP from the normal-velocity deviations (no cancellation).
            float dNP;
            if (fabsf(MN) < 1.f && fabsf(MP) < 1.f) {
                const float s2 = MN * MN + MN * MP + MP * MP;
                const float s4 = MN * MN * MN * MN + MN * MN * MN * MP + MN * MN * MP * MP + MN * MP * MP * MP +
                                 MP * MP * MP * MP;
                const float dd = 0.5f * (3.f - s2) + 2.f * alpha * (1.f - 2.f * s2 + s4);
                dNP = (unN - unP) * rSc * dd;
            } else {
                dNP = dN - dP;
            }
            const float phalf = a.p_inf * dNP + pNd * dN - pPd * dP;
            const float pu = -0.75f * p5pos32(MP, alpha) * p5neg32(MN, alpha) * (rhoPos + rhoNeg) * cavg * fa *
                             (unN - unP);
            pp += tha * (pu - 0.5f * phalf);
        }
        float mS = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) mS += mp[k] * S[k];
        rhs[0] = -mS;
#pragma unroll
        for (int k = 0; k < D; k++) rhs[k + 1] = -(phip * a.minf[k] + phi * mp[k] + pp * S[k]);
        rhs[D + 1] = -(phip * a.G_inf + phi * Gdev);
    }

    // viscous terms (cfd_v0.cpp:2682-2790) on velocity / temperature deviations
    if (VISC) {
    const float mu = a.mu;
    float dmag2 = 0.f;
#pragma unroll
    for (int k = 0; k < D; k++) dmag2 += d[k] * d[k];
    const float dmi = rsq32(dmag2);
    float dn[D], du[D][D], dT[D], tau[D][D], tmc[D][D];
    float duP[D][D], duN[D][D], sUP[D], sUN[D];
#pragma unroll
    for (int k = 0; k < D; k++) dn[k] = d[k] * dmi;
    const float cT = sp.Rp * a.R_inv, aT = sn.Rp * a.R_inv;
    if (ghost) {
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) du[i][j] = (sn.up[i] - sp.up[i]) * dn[j] * dmi;
            dT[i] = (aT - cT) * dn[i] * dmi;
        }
    } else {
#pragma unroll
        for (int i = 0; i < D; i++) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                duP[i][j] = gr[0][i * D + j];
                duN[i][j] = gr[1][i * D + j];
                du[i][j] = LERP32(w, duP[i][j], duN[i][j]);
            }
            dT[i] = LERP32(w, gr[0][R::kDT + i], gr[1][R::kDT + i]);
        }
    }
#pragma unroll
    for (int p = 0; p < D; p++) {
        tau[p][p] = 2.0f * du[p][p];
#pragma unroll
        for (int b1 = p + 1; b1 < D + p; b1++) {
            const int b = b1 % D;
            tau[p][b] = mu * (du[p][b] + du[b][p]);
            tau[p][p] -= du[b][b];
        }
        tau[p][p] *= 2.0f / 3.0f * mu;
    }
    if (ghost) {
        float diag = 0.f;
#pragma unroll
        for (int p = 0; p < D; p++) diag -= du[p][p];
        diag *= mu * (2.0f / 3.0f);
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
                tmc[i][j] = LERP32(w, a.tauMC[(i * D + j) * nt + P], a.tauMC[(i * D + j) * nt + N]);
    } else {
        // both cells' tensors as calc_VIS leaves them, from their dudx
        float tP[D][D], tN[D][D];
        vis_tensors32<D>(a, duP, sp.u, a.smag ? a.smag + P : nullptr, tP, sUP);
        vis_tensors32<D>(a, duN, sn.u, a.smag ? a.smag + N : nullptr, tN, sUN);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) tmc[i][j] = LERP32(w, tP[i][j], tN[i][j]);
    }
    const float Sd = dot32<D>(S, d);
    float K[D], dm2 = 0.f;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const float delta = div32(d[i] * Sm2, Sd);
        dm2 += delta * delta;
        K[i] = S[i] - delta;
    }
    const float dm = sqrt32(dm2);
    float sU[D];
    if (ghost) {
        float Uf[D];
#pragma unroll
        for (int i = 0; i < D; i++) Uf[i] = LERP32(w, sp.u[i], sn.u[i]);
#pragma unroll
        for (int i = 0; i < D; i++) sU[i] = dot32<D>(Uf, tau[i]);
    } else if constexpr (STORED) {
#pragma unroll
        for (int i = 0; i < D; i++) sU[i] = LERP32(w, a.sigmaU[i * nt + P], a.sigmaU[i * nt + N]);
    } else {
#pragma unroll
        for (int i = 0; i < D; i++) sU[i] = LERP32(w, sUP[i], sUN[i]);
    }
    const float kcond = a.cp * mu * a.pr_inv;
#pragma unroll
    for (int i = 0; i < D; i++)
        rhs[i + 1] += dot32<D>(tmc[i], S) + mu * (dm * (sn.up[i] - sp.up[i]) * dmi + dot32<D>(K, du[i]));
    rhs[D + 1] += dot32<D>(sU, S) + kcond * (dm * (aT - cT) * dmi + dot32<D>(K, dT));
    }
}

template <int D, int SOLVER, bool VISC, bool STORED, class GA>
__device__ __forceinline__ void face_flux32_rhs(const KArgs32 &a, const int f, const GA &ga, const int gP,
                                                const int gN, float (&rhs)[D + 2]) {
    const int P = ld_once<D, 3>(a.fP + f), N = ld_once<D, 3>(a.fN + f);
    float qP[D + 2], qN[D + 2];
    load_q32<D>(a, P, qP);
    load_q32<D>(a, N, qN);
    const Prim32<D> sp = prim32<D>(a, qP), sn = prim32<D>(a, qN);
    float S[D];
#pragma unroll
    for (int k = 0; k < D; k++) S[k] = ld_once<D, 3>(a.S + (size_t)f * D + k);
    face_flux32_states<D, SOLVER, VISC, STORED>(a, f, P, N, qP, qN, sp, sn, S, ld_once<D, 3>(a.w + f), ga, gP, gN,
                                                 rhs);
}

template <int D, int SOLVER, bool VISC, bool STORED>
__device__ __forceinline__ void face_flux32_body(const KArgs32 &a, const int f) {
    if (f >= a.nf) return;
    float rhs[D + 2];
    face_flux32_rhs<D, SOLVER, VISC, STORED>(a, f, GradGlobal32<D, SOLVER == 2>{a}, a.fP[f], a.fN[f], rhs);
    st4_once<D, 3>(a.frhs + 4 * (size_t)f, rhs[0], rhs[1], rhs[2], rhs[3]);
    if (D == 3) st_once<D, 3>(a.frhs + (size_t)4 * a.nf + f, rhs[D + 1]);
}

// ---------------------------------------------------------------------------
// ordered gather of face fluxes, sponge, RK update (deviations)
// ---------------------------------------------------------------------------
// What the ordered gather of a cell (cell_update32_cell) asks of its source
// of face data, for the cell's j-th gather entry e (j < kMaxFaces, known at
// compile time wherever the gather calls these):
//   key(e, j, c)       an entry key, >= 0 where the cell is the face's P;
//   ghost(key)         the boundary ghost behind a P entry, else -1;
//   get(e, j, key, r)  the face's flux;
//   area(e, j, c, S)   the face's area vector, negated on N entries.
// kUnroll entries have their loads issued together.  FaceList32 reads the
// gather list (key = the signed face index with the ghost flag);
// FluxGlobal32 adds the stored face fluxes.
template <int D>
struct FaceList32 {
    const KArgs32 &a;
    static constexpr int kUnroll = UpdateUnroll32<D>::value;
    static constexpr bool kAllSlots = false;
    static constexpr bool kSlotsFirst = false;
    __device__ __forceinline__ static bool valid(int key, int c) { return (void)key, (void)c, true; }
    __device__ __forceinline__ void begin_cell(int c) const { (void)c; }
    __device__ __forceinline__ int key(int e, int j, int c) const {
        (void)j;
        (void)c;
        return a.g_ent[e];
    }
    __device__ __forceinline__ int ghost(int key) const {
        return (key >= 0 && (key & kGhostFaceFlag)) ? a.fN[key & kFaceMask] - a.nc : -1;
    }
    __device__ __forceinline__ static int face(int key) { return key >= 0 ? (key & kFaceMask) : ~key; }
    __device__ __forceinline__ void area(int e, int j, int c, float (&S)[D]) const {
        const int key = this->key(e, j, c);
        load_S32<D>(a.S, face(key), S);
#pragma unroll
        for (int k = 0; k < D; k++) S[k] = key >= 0 ? S[k] : -S[k];
    }
};

template <int D>
struct FluxGlobal32 : FaceList32<D> {
    using FaceList32<D>::a;
    using FaceList32<D>::face;
    // the flux of gather entry e
    __device__ __forceinline__ void get(int e, int j, int key, float (&r)[D + 2]) const {
        (void)e;
        (void)j;
        const int f = face(key);
        const float4 v = reinterpret_cast<const float4 *>(a.frhs)[f];
        r[0] = v.x;
        r[1] = v.y;
        r[2] = v.z;
        r[3] = v.w;
        if (D == 3) r[D + 1] = a.frhs[(size_t)4 * a.nf + f];
    }
};
// ordered gather of one cell's face fluxes, sponge and RK update from the
// state in a.q, written to qout (a.q itself for the in-place update) with
// the cell's boundary ghosts refreshed there; res receives the residual, q
// the updated state
template <int D, bool RES, class FA>
__device__ __forceinline__ void cell_update32_cell(const KArgs32 &a, const FA &fa, float A, float B, int zero_dq,
                                                   const int c, float *qout, float (&res)[D + 2],
                                                   float (&q)[D + 2], float &vinv) {
    const float dt = (float)(*a.dt);
    vinv = ld_once<D, 3>(a.vinv + c);
    float dq[D + 2];
#pragma unroll
    for (int k = 0; k < D + 2; k++) {
        dq[k] = zero_dq ? 0.f : ld_once<D, 2>(a.dq + (size_t)k * a.nc + c) * A;
        res[k] = 0.f;
    }
    // kAllSlots: every entry is in the face slots (placeholders past the
    // last), so the gather list's offsets are not read; the group count is
    // then kept opaque to the compiler, which would otherwise set up every
    // group's addresses at once and spill
    int e0, ne;
    if constexpr (FA::kAllSlots) {
        e0 = 0;
        ne = kMaxFaces;
        asm volatile("" : "+s"(ne));
    } else {
        e0 = a.g_off[c];
        ne = a.g_off[c + 1] - e0;
    }
    unsigned gmask = 0;  // bit j: entry j leads to a boundary ghost
    // entry j's flux added (valid: j is one of the cell's entries; an
    // invalid one is selected away, not branched around, so that the loads
    // of a group stay together)
    auto add = [&](const int j, const bool valid, const int key, const float (&fr)[D + 2]) {
        gmask |= valid && fa.ghost(key) >= 0 ? 1u << j : 0u;
        const float sg = key >= 0 ? 1.f : -1.f;
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            const float r = valid ? sg * fr[k] : 0.f;
            if (RES) res[k] = valid ? res[k] + r : res[k];
            dq[k] = valid ? dq[k] + dt * r * vinv : dq[k];
        }
    };
    // the faces in gather order, kUnroll at a time with their loads issued
    // together (a cell has at most kMaxFaces).  Within a group no branch: an
    // entry past the cell's last reads the last entry again (list) or the
    // slot's placeholder, and is not added.
    constexpr int U = FA::kUnroll;
    fa.begin_cell(c);
#pragma unroll
    for (int j0 = 0; j0 < kMaxFaces; j0 += U) {
        // every cell has a first group: where its loads do not address
        // through the gather list, they need not wait for ne
        if ((j0 == 0 && FA::kSlotsFirst) || j0 < ne) {
            int key[U];
            float fr[U][D + 2];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int j = j0 + u;
                if (j < kMaxFaces) {
                    const int e = e0 + (j < ne ? j : ne - 1);
                    key[u] = fa.key(e, j, c);
                    fa.get(e, j, key[u], fr[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j0 + u < kMaxFaces) add(j0 + u, FA::kAllSlots ? fa.valid(key[u], c) : j0 + u < ne, key[u], fr[u]);
        }
    }
    load_q32<D>(a, c, q);
    // sponge: sigma (q_inf - q) = -sigma q' (no sigma array, and no term,
    // where the case has no sponge layer and every sigma is zero)
    const float sgm = a.sigma ? ld_once<D, 3>(a.sigma + c) : 0.f;
#pragma unroll
    for (int k = 0; k < D + 2; k++) {
        if (a.sigma) dq[k] -= dt * sgm * q[k];
        q[k] += B * dq[k];
        st_once<D, 2>(a.dq + (size_t)k * a.nc + c, dq[k]);
    }
    store_q32<D>(qout, a.ntot, c, q);
    // refresh this cell's boundary ghosts for the next stage
    for (unsigned m = gmask; m; m &= m - 1) {
        const int j = __builtin_ctz(m);
        const int g = fa.ghost(fa.key(e0 + j, j, c));
        float qg[D + 2];
        if (!ghost_state32<D>(a, a.ghost_role[g], q, qg)) {
            if (qout == a.q) continue;
            load_q32<D>(a, a.nc + g, qg);  // an unset patch's ghost keeps its state
        }
        store_q32<D>(qout, a.ntot, a.nc + g, qg);
    }
}

template <int D, int MODE, class FA>
__device__ __forceinline__ double cfl_cell32(const KArgs32 &a, const FA &fa, const int c, const float (&qc)[D + 2],
                                             float vinv, double param);

// every thread of the block must call this (block-wide reductions when RES
// or CFL).  CFL >= 0 (the last stage of a step): the step's compute_dt (0) /
// compute_cfl (1) of the updated cells, per-block extrema into a.partial.
template <int D, bool RES, int CFL, class FA>
__device__ __forceinline__ void cell_update32_body(const KArgs32 &a, const FA &fa, float *qout, float A, float B,
                                                   int zero_dq, const int c, const int lblk, double cfl_param) {
    static_assert(!(RES && CFL >= 0), "the residual and the CFL share a.partial");
    __shared__ double red[256 / 64][D + 2];
    double rsq[D + 2];
    double cv = CFL == 0 ? 100.0 : 0.0;
#pragma unroll
    for (int k = 0; k < D + 2; k++) rsq[k] = 0.0;
    if (c < a.nc) {
        float res[D + 2], q[D + 2], vinv;
        cell_update32_cell<D, RES>(a, fa, A, B, zero_dq, c, qout, res, q, vinv);
        if constexpr (CFL >= 0) cv = cfl_cell32<D, CFL>(a, fa, c, q, vinv, cfl_param);
        if (RES) {
#pragma unroll
            for (int k = 0; k < D + 2; k++) {
                a.res_cell[(size_t)k * a.nc + c] = res[k];
                rsq[k] = (double)res[k] * (double)res[k];
            }
        }
    }
    if (RES) {
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
                a.partial[(size_t)lblk * (D + 2) + k] = s;
            }
        }
    }
    if constexpr (CFL >= 0) block_extremum<CFL>(cv, a.partial + lblk);
}

// compute_dt / compute_cfl (cfd_v0.cpp:2887-2942) of one cell from its FP32
// deviations q: the face terms |m.S| (compute_cfl) or |m.S|/|S| (compute_dt)
// summed in FP32 over the cell's faces in gather order (the update's own
// list, whose entries and area vectors it has just read; the FP64 path keeps
// the reference's slot order), then dt_c (MODE 0) or cfl_c (MODE 1) in FP64,
// combined with the block's starting value so a NaN never wins; vinv:
// the cell's 1/V, as the update read it
template <int D, int MODE, class FA>
__device__ __forceinline__ double cfl_cell32(const KArgs32 &a, const FA &fa, const int c, const float (&qc)[D + 2],
                                             float vinv_c, double param) {
    float m[D];
#pragma unroll
    for (int k = 0; k < D; k++) m[k] = a.minf[k] + qc[k + 1];
    // kAllSlots: every entry in the slots, a placeholder known by its zero
    // area vector (a face has a nonzero one)
    const int e0 = FA::kAllSlots ? 0 : a.g_off[c];
    const int ne = FA::kAllSlots ? kMaxFaces : a.g_off[c + 1] - e0;
    float sum = 0.f;
#pragma unroll
    for (int j = 0; j < kMaxFaces; j++) {
        float S[D];
        fa.area(e0 + (j < ne ? j : ne - 1), j, c, S);
        float fl = 0.f, s2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; k++) fl += m[k] * S[k];
#pragma unroll
        for (int k = 0; k < D; k++) s2 += S[k] * S[k];
        if (MODE == 0) fl *= rsq32(s2);
        sum = (FA::kAllSlots ? s2 > 0.f : j < ne) ? sum + fabsf(fl) : sum;
    }
    const double rho = a.rho_inf_d + (double)qc[0];
    const double vinv = (double)vinv_c;
    double v = MODE == 0 ? 100.0 : 0.0;
    if (MODE == 0) {
        const double dtc = param * rho * 2.0 / (vinv * (double)sum);
        if (v > dtc) v = dtc;
    } else {
        const double cfl = (double)sum * ((1.0 / rho) * vinv * param);
        if (v < cfl) v = cfl;
    }
    return v;
}

// compute_dt / compute_cfl as their own launch (every thread of the block
// must call this)
template <int D, int MODE>
__device__ __forceinline__ void cfl_dt_cells32_body(const KArgs32 &a, double param, const int c, const int lblk) {
    double v = MODE == 0 ? 100.0 : 0.0;
    if (c < a.nc) {
        float q[D + 2];
        load_q32<D>(a, c, q);
        v = cfl_cell32<D, MODE>(a, FaceList32<D>{a}, c, q, a.vinv[c], param);
    }
    block_extremum<MODE>(v, a.partial + lblk);
}

// FP64 conservation integrals sum_c V_c q_c (per block partials)
template <int D>
__global__ void __launch_bounds__(256) integrals32(KArgs32 a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double red[256 / 64][D + 2];
    double v[D + 2];
#pragma unroll
    for (int k = 0; k < D + 2; k++) v[k] = 0.0;
    if (c < a.nc) {
        const double vol = a.vol_d[c];
        v[0] = vol * (a.rho_inf_d + (double)a.q[q32_index(0, c, a.ntot)]);
#pragma unroll
        for (int k = 0; k < D; k++) v[k + 1] = vol * (a.m_inf_d[k] + (double)a.q[q32_index(k + 1, c, a.ntot)]);
        v[D + 1] = vol * (a.rhoE_inf_d + (double)a.q[q32_index(D + 1, c, a.ntot)]);
    }
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

template <int D>
__global__ void __launch_bounds__(256) count_nonfinite32(const float *q, int nc, int ntot, unsigned long long *cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int bad = 0;
#pragma unroll
    for (int k = 0; k < D + 2; k++) bad += !isfinite(q[q32_index(k, c, ntot)]);
    if (bad) atomicAdd(cnt, (unsigned long long)bad);
}

// ---------------------------------------------------------------------------
// single-solver kernels
// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256) bc_fill32(KArgs32 a) {
    bc_fill32_body<D>(a, LFM_TID(blockIdx.x));
}
// fin_nblk > 0: one block more, the last, reduces the previous step's
// fin_nblk per-block CFL maxima in a.partial into *a.d_cfl (as the one-launch
// stage does)
template <int D, int GRAD, bool VISC, bool RG, bool SL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(D == 3 && SL && !RG ? 5 : 1)))
grad_vis32(KArgs32 a, int fin_nblk) {
    int nb = (int)gridDim.x;
    if (fin_nblk > 0) {
        nb -= 1;
        if ((int)blockIdx.x == nb) {
            cfl_dt_final_block<1, 256>(a.partial, fin_nblk, a.d_cfl, nullptr);
            return;
        }
    }
    grad_vis32_body<D, GRAD, VISC, RG, SL>(a, xcd_block((int)blockIdx.x, nb) * (int)blockDim.x + (int)threadIdx.x);
}
// 3D M1 / M2 faces: 5 waves per SIMD fit in 96 VGPRs without spills (97
// would allow 4): O-grid face kernel 1.107 -> 1.033 ms per step.  Forcing
// more waves on the update and gradient kernels was measured slower.
template <int D, int SOLVER, bool VISC, bool STORED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(D == 3 && SOLVER < 2 ? 5 : 1)))
face_flux32(KArgs32 a) {
    face_flux32_body<D, SOLVER, VISC, STORED>(a, LFM_TID(blockIdx.x));
}
template <int D, bool RES, int CFL>
__global__ void __launch_bounds__(256) cell_update32(KArgs32 a, float A, float B, int zero_dq, double cfl_param) {
    cell_update32_body<D, RES, CFL>(a, FluxGlobal32<D>{a}, a.q, A, B, zero_dq, LFM_TID(blockIdx.x), LFM_BLK(),
                                    cfl_param);
}
// Where a fused stage takes a face's flux: computed where it is gathered, by
// the face kernel's own function, from the stage's input state a.q
// (inviscid M1 / M2: no gradients), with the gathering cell's state and
// primitives loaded and formed once per cell.  The face data come from the
// cell's slots (a.sl_*: streamed, one per gather entry) for its first
// GatherUnroll entries (all of a 3D cell's) and from the face arrays beyond
// them.  Each interior
// face is computed by both of its cells, to the same bits as the face
// kernel's (the same function of the same values).
template <int D, int SOLVER>
struct FluxSlots32 {
    const KArgs32 &a;
    // LDS for the slots' area vectors, oriented out of the cell, kept for
    // the CFL/dt sums after the update ([U][D][256] per block; nullptr: not
    // kept)
    float *keep_S;
    static constexpr int U = GatherUnroll<D>::value;  // the slots per cell
    // 3D: two faces' loads issued together, which leaves 63 VGPRs and 8
    // waves per SIMD (all six at 4 waves, or three, were slower)
    static constexpr int kUnroll = D == 3 ? 2 : U;
    // a 3D cell's entries all fit the slots (kMaxFaces of them)
    static constexpr bool kAllSlots = U >= kMaxFaces;
    // the first kUnroll entries are slots, addressed by the cell alone, so
    // their loads can issue before the entry count arrives (2D: 0.0271 ->
    // 0.0263 ms per 65 k-cell step; the 3D box showed no gain)
    static constexpr bool kSlotsFirst = D == 2;
    // a slot past the cell's last entry holds the cell itself
    __device__ __forceinline__ static bool valid(int key, int c) { return key != c; }
    mutable float qc[D + 2];
    mutable Prim32<D> pc;
    mutable int c_ = 0;
    __device__ __forceinline__ void begin_cell(int c) const {
        c_ = c;
        load_q32<D>(a, c, qc);
        pc = prim32<D>(a, qc);
    }
    // key: the face's other cell, ~other where the cell is the face's N;
    // from the slots up to U, from the face arrays beyond
    __device__ __forceinline__ int key(int e, int j, int c) const {
        if (j < U) return ld_once<D, 1>(a.sl_nb + slot_at<D, 1>(a.nc, j, 0, c));
        const int ent = a.g_ent[e];
        return ent >= 0 ? a.fN[ent & kFaceMask] : ~a.fP[~ent];
    }
    // boundary ghosts sit after the cells (no halo ghosts on this path)
    __device__ __forceinline__ int ghost(int key) const { return key >= a.nc ? key - a.nc : -1; }
    __device__ __forceinline__ void flux(int key, const float (&S)[D], float w, float (&r)[D + 2]) const {
        const bool isP = key >= 0;
        const int o = isP ? key : ~key;
        float qo[D + 2], qP[D + 2], qN[D + 2];
        load_q32<D>(a, o, qo);
        const Prim32<D> po = prim32<D>(a, qo);
#pragma unroll
        for (int k = 0; k < D + 2; k++) {
            qP[k] = isP ? qc[k] : qo[k];
            qN[k] = isP ? qo[k] : qc[k];
        }
        const Prim32<D> sp = isP ? pc : po, sn = isP ? po : pc;
        face_flux32_states<D, SOLVER, false, false>(a, 0, 0, 0, qP, qN, sp, sn, S, w, GradGlobal32<D, false>{a}, 0,
                                                     0, r);
    }
    // area vector and weight of entry j
    __device__ __forceinline__ void data(int e, int j, float (&S)[D], float &w) const {
        if (j < U) {
#pragma unroll
            for (int k = 0; k < D; k++) S[k] = ld_once<D, 1>(a.sl_S + slot_at<D, D>(a.nc, j, k, c_));
            w = ld_once<D, 1>(a.sl_w + slot_at<D, 1>(a.nc, j, 0, c_));
            return;
        }
        const int ent = a.g_ent[e];
        const int f = ent >= 0 ? (ent & kFaceMask) : ~ent;
        load_S32<D>(a.S, f, S);
        w = a.w[f];
    }
    __device__ __forceinline__ void get(int e, int j, int key, float (&r)[D + 2]) const {
        float S[D], w;
        data(e, j, S, w);
        if (keep_S && j < U) {
#pragma unroll
            for (int k = 0; k < D; k++) keep_S[(j * D + k) * 256 + threadIdx.x] = key >= 0 ? S[k] : -S[k];
        }
        flux(key, S, w, r);
    }
    __device__ __forceinline__ void area(int e, int j, int c, float (&S)[D]) const {
        if (keep_S && j < U) {
#pragma unroll
            for (int k = 0; k < D; k++) S[k] = keep_S[(j * D + k) * 256 + threadIdx.x];
            return;
        }
        float w;
        data(e, j, S, w);
        const bool isP = key(e, j, c) >= 0;
#pragma unroll
        for (int k = 0; k < D; k++) S[k] = isP ? S[k] : -S[k];
    }
};

// the one-launch stage's face slots from the gather list: slot j of cell c
// holds its j-th gather entry's other cell (~other on N entries), area
// vector and weight
template <int D>
__global__ void __launch_bounds__(256) build_slots32(KArgs32 a, int32_t *nb, float *S, float *w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    const int e0 = a.g_off[c], ne = a.g_off[c + 1] - e0;
    constexpr int U = GatherUnroll<D>::value;
    for (int j = 0; j < U; j++) {
        if (j >= ne) {
            // placeholder past the cell's last entry: the cell itself, zero
            // area (read by unbranched gathers, never added)
            nb[slot_at<D, 1>(a.nc, j, 0, c)] = c;
#pragma unroll
            for (int k = 0; k < D; k++) S[slot_at<D, D>(a.nc, j, k, c)] = 0.f;
            w[slot_at<D, 1>(a.nc, j, 0, c)] = 0.5f;
            continue;
        }
        const int ent = a.g_ent[e0 + j];
        const int f = ent >= 0 ? (ent & kFaceMask) : ~ent;
        nb[slot_at<D, 1>(a.nc, j, 0, c)] = ent >= 0 ? a.fN[f] : ~a.fP[f];
#pragma unroll
        for (int k = 0; k < D; k++) S[slot_at<D, D>(a.nc, j, k, c)] = a.S[(size_t)f * D + k];
        w[slot_at<D, 1>(a.nc, j, 0, c)] = a.w[f];
    }
}

// One RK stage in one launch for inviscid M1 / M2: the face fluxes, the
// ordered gather, sponge and update of cell_update32, reading the state in
// a.q and writing the new state and boundary ghosts to qout (another buffer).
// 3D: at least 7 waves per SIMD (6 for the last stage, whose CFL/dt sums
// keep more live; both compile to 8), with two faces' loads in flight per
// wave (kUnroll).
// fin_nblk > 0 (a stage without CFL/dt sums): the grid has one block more,
// the last, which reduces the fin_nblk per-block CFL maxima that an earlier
// launch left in a.partial into *a.d_cfl (the final reduction of the
// previous step, whose result no stage of this step reads).
template <int D, int SOLVER, int CFL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(D == 3 ? (CFL < 0 ? 7 : 6) : 1)))
stage_fused32(KArgs32 a, float *qout, float A, float B, int zero_dq, double cfl_param, int fin_nblk) {
    // the last stage keeps its cells' area vectors in LDS for the CFL/dt sums
    __shared__ float keep[CFL >= 0 ? GatherUnroll<D>::value * D * 256 : 1];
    int nb = (int)gridDim.x;
    if constexpr (CFL < 0) {
        if (fin_nblk > 0) {
            nb -= 1;
            if ((int)blockIdx.x == nb) {
                cfl_dt_final_block<1, 256>(a.partial, fin_nblk, a.d_cfl, nullptr);
                return;
            }
        }
    } else {
        (void)fin_nblk;
    }
    const int blk = xcd_block((int)blockIdx.x, nb);
    cell_update32_body<D, false, CFL>(a, FluxSlots32<D, SOLVER>{a, CFL >= 0 ? keep : nullptr}, qout, A, B, zero_dq,
                                      blk * (int)blockDim.x + (int)threadIdx.x, blk, cfl_param);
}
template <int D, int MODE>
__global__ void __launch_bounds__(256) cfl_dt_cells32(KArgs32 a, double param) {
    cfl_dt_cells32_body<D, MODE>(a, param, LFM_TID(blockIdx.x), LFM_BLK());
}

// tauMC and sigmaU of every cell from the stored dudx (field read-back of a
// solver that does not store them): [D*D][nc] and [D][nc] deviations-free
// tensors (they do not depend on the free stream)
template <int D>
__global__ void __launch_bounds__(256) vis_tensor_field32(KArgs32 a, float *tmc_out, float *sU_out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    float du[D][D], q[D + 2], tmc[D][D], sU[D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) du[i][j] = a.dudx[(size_t)c * a.gstride + i * D + j];
    load_q32<D>(a, c, q);
    const Prim32<D> s = prim32<D>(a, q);
    vis_tensors32<D>(a, du, s.u, a.smag ? a.smag + c : nullptr, tmc, sU);
#pragma unroll
    for (int i = 0; i < D; i++) {
        sU_out[(size_t)i * a.nc + c] = sU[i];
#pragma unroll
        for (int j = 0; j < D; j++) tmc_out[(size_t)(i * D + j) * a.nc + c] = tmc[i][j];
    }
}

// ---------------------------------------------------------------------------
// batched kernels: one launch covers several resident solvers of one scheme
// class.  Block b belongs to solver bmap[b] (sorted) and is its local block
// b - bfirst[solver]; each solver computes exactly what its single launch
// computes.
// ---------------------------------------------------------------------------
struct BatchMap {
    const int32_t *req;    // [nblocks] solver of each block
    const int32_t *first;  // [nsolvers] first block of each solver
};

__device__ __forceinline__ void batch_slot(const BatchMap &m, int &r, int &lblk) {
    const int b = LFM_BLK();
    r = m.req[b];
    lblk = b - m.first[r];
}

template <int D, int GRAD, bool VISC, bool RG>
__global__ void __launch_bounds__(256) grad_vis32_b(const KArgs32 *args, BatchMap m) {
    int r, lb;
    batch_slot(m, r, lb);
    grad_vis32_body<D, GRAD, VISC, RG>(args[r], lb * 256 + threadIdx.x);
}
// (batch members are undecomposed: their tensors are never stored)
template <int D, int SOLVER, bool VISC>
__global__ void __launch_bounds__(256) face_flux32_b(const KArgs32 *args, BatchMap m) {
    int r, lb;
    batch_slot(m, r, lb);
    face_flux32_body<D, SOLVER, VISC, false>(args[r], lb * 256 + threadIdx.x);
}
// CFL 1: the last stage, with compute_cfl at each solver's own deltaT
template <int D, int CFL>
__global__ void __launch_bounds__(256) cell_update32_b(const KArgs32 *args, BatchMap m, float A, float B,
                                                       int zero_dq) {
    int r, lb;
    batch_slot(m, r, lb);
    const KArgs32 &a = args[r];
    cell_update32_body<D, false, CFL>(a, FluxGlobal32<D>{a}, a.q, A, B, zero_dq, lb * 256 + threadIdx.x, lb,
                                      CFL >= 0 ? *a.dt : 0.0);
}
// one-launch stage over a batch (inviscid M1 / M2), each member from one
// of its state buffers into the other: flip 0 from a.q into a.q_alt, flip 1
// back
template <int D, int SOLVER, int CFL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(D == 3 ? (CFL < 0 ? 7 : 6) : 1)))
stage_fused32_b(const KArgs32 *args, BatchMap m, float A, float B, int zero_dq, int flip) {
    __shared__ float keep[CFL >= 0 ? GatherUnroll<D>::value * D * 256 : 1];
    int r, lb;
    batch_slot(m, r, lb);
    KArgs32 a = args[r];
    float *qout = flip ? a.q : a.q_alt;
    if (flip) a.q = a.q_alt;
    cell_update32_body<D, false, CFL>(a, FluxSlots32<D, SOLVER>{a, CFL >= 0 ? keep : nullptr}, qout, A, B, zero_dq,
                                      lb * 256 + threadIdx.x, lb, CFL >= 0 ? *a.dt : 0.0);
}
// one block per solver: max of its partials into its CFL scalar
static __global__ void __launch_bounds__(256) cfl_final32_b(const KArgs32 *args) {
    __shared__ double sh[256];
    const KArgs32 &a = args[blockIdx.x];
    double v = 0.0;
    // min / max are exact in any order: keep several loads in flight
#pragma unroll 8
    for (int i = threadIdx.x; i < a.nblk_cells; i += blockDim.x) {
        const double o = a.partial[i];
        if (v < o) v = o;
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st && sh[threadIdx.x] < sh[threadIdx.x + st]) sh[threadIdx.x] = sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *a.d_cfl = sh[0];
}

// launchers (defined in solver_fp32p.hip)
template <int D>
void launch_bc32(Solver *s);
template <int D>
void launch_gradvis32(Solver *s, bool full, bool fin = false);
template <int D>
void launch_flux32(Solver *s);
template <int D>
void launch_update32(Solver *s, double A, double B, bool zero_dq, bool res, bool accumulate_res,
                     const Solver::CflJob *cfl);
template <int D>
void launch_cfl_dt32(Solver *s, int mode, double param, double *out, double *dt_dst);
template <int D>
void launch_build_slots32(Solver *s, int32_t *nb, float *S, float *w);
template <int D>
void launch_stage_fused32(Solver *s, double A, double B, bool zero_dq, const float *qin, float *qout,
                          const Solver::CflJob *cfl, bool fin);

}  // namespace lfm
