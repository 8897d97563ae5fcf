// solver.hip -- device state, launches and the C ABI (include/lfm_hip.h).
//
// One lfm_solver holds one partition (both reference submeshes) resident in
// HBM in SoA layout and drives the stage kernels on its own HIP stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kernels_fp64.hpp"
#include "kernels_fp32p.hpp"
#include "lfm_hip.h"
#include "mesh_builder.hpp"
#include "solver_state.hpp"

namespace lfm {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

std::mutex &device_setup_mutex() {
    static std::mutex m;
    return m;
}

static constexpr int kBlock = 256;
static inline int nblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
static int dev_upload(Solver *s, T **dst, const T *src, size_t n) {
    if (n == 0) {
        *dst = nullptr;
        return 0;
    }
    HIP_TRY(hipMalloc((void **)dst, n * sizeof(T)));
    s->allocs.push_back((void *)*dst);
    s->device_bytes += (int64_t)(n * sizeof(T));
    if (src) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
static int dev_zeros(Solver *s, T **dst, size_t n) {
    if (n == 0) {
        *dst = nullptr;
        return 0;
    }
    HIP_TRY(hipMalloc((void **)dst, n * sizeof(T)));
    s->allocs.push_back((void *)*dst);
    s->device_bytes += (int64_t)(n * sizeof(T));
    HIP_TRY(hipMemset(*dst, 0, n * sizeof(T)));
    return 0;
}

Solver::~Solver() {
    for (void *p : allocs) (void)hipFree(p);
    if (k32) free_k32(k32);
    free_post(post);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    if (h_scalar) (void)hipHostFree(h_scalar);
}

// ---------------------------------------------------------------------------
// creation
// ---------------------------------------------------------------------------
static int create_fp64(Solver *s) {
    const SolverLayout &L = s->L;
    const int D = L.D;
    const size_t nc = L.nc, ntot = (size_t)L.nc + L.ng, nf = L.nf;
    KArgs &a = s->k64;
    std::memset(&a, 0, sizeof(a));
    a.nc = L.nc;
    a.nsm0 = L.nsm0;
    a.ng = L.ng;
    a.nf = L.nf;
    a.ng_bnd = L.ng_bnd;
    a.ntot = (int)ntot;
    int rc = 0;
    rc |= dev_upload(s, (int32_t **)&a.fP, L.fP.data(), nf);
    rc |= dev_upload(s, (int32_t **)&a.fN, L.fN.data(), nf);
    {
        // S as one record per face ([nf][D])
        std::vector<double> Sr((size_t)D * nf);
        for (size_t f = 0; f < nf; f++)
            for (int k = 0; k < D; k++) Sr[f * D + k] = L.S[(size_t)k * nf + f];
        rc |= dev_upload(s, (double **)&a.S, Sr.data(), Sr.size());
    }
    rc |= dev_upload(s, (double **)&a.d, L.d.data(), (size_t)D * nf);
    rc |= dev_upload(s, (double **)&a.w, L.w.data(), nf);
    rc |= dev_upload(s, (int32_t **)&a.g_off, L.g_off.data(), nc + 1);
    rc |= dev_upload(s, (int32_t **)&a.g_ent, L.g_ent.data(), L.g_ent.size());
    rc |= dev_upload(s, (int32_t **)&a.s_off, L.s_off.data(), nc + 1);
    rc |= dev_upload(s, (int32_t **)&a.s_ent, L.s_ent.data(), L.s_ent.size());
    rc |= dev_upload(s, (int32_t **)&a.ghost_cell, L.ghost_cell.data(), L.ng);
    rc |= dev_upload(s, (int32_t **)&a.ghost_role, L.ghost_role.data(), L.ng);
    rc |= dev_upload(s, (double **)&a.vinv, L.vinv.data(), nc);
    rc |= dev_upload(s, (double **)&a.sigma, L.sigma.data(), nc);
    if (!L.smag.empty()) rc |= dev_upload(s, (double **)&a.smag, L.smag.data(), nc);
    rc |= dev_zeros(s, &a.q, (size_t)(D + 2) * ntot);
    rc |= dev_zeros(s, &a.dq, (size_t)(D + 2) * nc);
    rc |= dev_zeros(s, &a.res_cell, (size_t)(D + 2) * nc);
    // gradient records (GradRec64): dudx, dTdx and (M2AUSM) the rho, p and U
    // gradients; solver 2 without minmodExists reads their never-computed
    // (zero) slots
    const bool rg = s->prm.solver == 2;
    a.gstride = grad_record_stride64(D, rg);
    rc |= dev_zeros(s, &a.dudx, (size_t)a.gstride * ntot);
    // tauMC / sigmaU are kept only for the halo payload of a partition; the
    // face kernel otherwise recomputes them from dudx
    if (!L.halo.empty()) {
        rc |= dev_zeros(s, &a.tauMC, (size_t)D * D * ntot);
        rc |= dev_zeros(s, &a.sigmaU, (size_t)D * ntot);
    }
    if (rc) return LFM_EDEVICE;
    a.dTdx = a.dudx + D * D;
    if (rg) {
        a.rho_grad = a.dudx + D * D + D;
        a.p_grad = a.dudx + D * D + 2 * D;
        a.U_grad = a.dudx + D * D + 3 * D;
    }
    rc |= dev_zeros(s, &a.frhs, (size_t)(D + 2) * nf);
    const size_t npart = (size_t)nblocks((long)std::max(nc, nf)) * (D + 2) + 16;
    rc |= dev_zeros(s, &a.partial, npart);
    rc |= dev_zeros(s, &s->d_scalar, 16);
    if (rc) return LFM_EDEVICE;
    a.dt = s->d_scalar + 0;
    s->d_res = s->d_scalar + 2;  // D+2 <= 5 slots: [2, 7)
    s->d_cfl = s->d_scalar + 8;
    s->d_dtmin = s->d_scalar + 9;
    const Gas &g = s->gas;
    a.gamma = g.gamma;
    a.gm1 = g.gm1;
    a.R_inv = g.R_inv;
    a.mu = g.mu;
    a.cp = g.cp;
    a.pr_inv = g.pr_inv;
    a.p_inf = g.p_inf;
    a.T_inf = g.T_inf;
    a.rho_inf = g.rho_inf;
    a.E_inf = g.E_inf;
    for (int k = 0; k < 3; k++) a.U_inf[k] = g.U_inf[k];
    // initial state into q (cells part, as cell records) and deltaT
    {
        std::vector<double> q0((size_t)(D + 2) * ntot, 0.0);
        for (int k = 0; k < D + 2; k++)
            for (size_t c = 0; c < nc; c++) q0[q64_index(k, c, ntot)] = L.q0[(size_t)k * nc + c];
        HIP_TRY(hipMemcpy(a.q, q0.data(), q0.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(s->d_scalar, &s->dt, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// ---------------------------------------------------------------------------
// FP64 launches
// ---------------------------------------------------------------------------
template <int D>
static void launch_bc64(Solver *s) {
    if (s->L.ng) hipLaunchKernelGGL(bc_fill<D>, dim3(nblocks(s->L.ng)), dim3(kBlock), 0, s->stream, s->k64);
}

template <int D, bool V>
static void launch_gradvis64_v(Solver *s) {
    const int nb = nblocks(s->L.nc);
    if (s->prm.solver == 2 && s->prm.minmod)
        hipLaunchKernelGGL((grad_vis<D, 2, V, true>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64);
    else if (s->prm.solver == 2)
        hipLaunchKernelGGL((grad_vis<D, 0, V, true>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64);
    else
        hipLaunchKernelGGL((grad_vis<D, 0, V, false>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64);
}

// full: every tensor calc_VIS forms (the per-stage entry point); otherwise
// only what the step consumes (nothing for solvers 0/1 when mu = 0)
template <int D>
static void launch_gradvis64(Solver *s, bool full) {
    if (full || s->visc)
        launch_gradvis64_v<D, true>(s);
    else if (s->prm.solver == 2)
        launch_gradvis64_v<D, false>(s);
}

template <int D, bool V, bool ST>
static void launch_flux64_v(Solver *s) {
    const int nb = nblocks(s->L.nf);
    switch (s->prm.solver) {
    case 0: hipLaunchKernelGGL((face_flux<D, 0, V, ST>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64); break;
    case 1: hipLaunchKernelGGL((face_flux<D, 1, V, ST>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64); break;
    default: hipLaunchKernelGGL((face_flux<D, 2, V, ST>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64); break;
    }
}

template <int D>
static void launch_flux64(Solver *s) {
    if (s->visc)
        s->k64.tauMC ? launch_flux64_v<D, true, true>(s) : launch_flux64_v<D, true, false>(s);
    else
        launch_flux64_v<D, false, false>(s);
}

template <int D>
static void launch_update64(Solver *s, double A, double B, bool zero_dq, bool res, bool accumulate_res,
                            const Solver::CflJob *cfl) {
    const int nb = nblocks(s->L.nc);
    if (res) {
        hipLaunchKernelGGL((cell_update<D, true, -1>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64, A, B,
                           (int)zero_dq, 0.0);
        hipLaunchKernelGGL(reduce_sum, dim3(1), dim3(kBlock), 0, s->stream, s->k64.partial, nb, D + 2, s->d_res,
                           (int)accumulate_res);
    } else if (cfl && cfl->mode == 0) {
        hipLaunchKernelGGL((cell_update<D, false, 0>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64, A, B,
                           (int)zero_dq, cfl->param);
    } else if (cfl) {
        hipLaunchKernelGGL((cell_update<D, false, 1>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64, A, B,
                           (int)zero_dq, cfl->param);
    } else {
        hipLaunchKernelGGL((cell_update<D, false, -1>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64, A, B,
                           (int)zero_dq, 0.0);
    }
}

template <int D>
static void launch_cfl_dt64(Solver *s, int mode, double param, double *out, double *dt_dst) {
    const int nb = nblocks(s->L.nc);
    if (mode == 0) {
        hipLaunchKernelGGL((cfl_dt_cells<D, 0>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64, param);
        hipLaunchKernelGGL(cfl_dt_final<0>, dim3(1), dim3(kFinalBlock), 0, s->stream, s->k64.partial, nb, out, dt_dst);
    } else {
        hipLaunchKernelGGL((cfl_dt_cells<D, 1>), dim3(nb), dim3(kBlock), 0, s->stream, s->k64, param);
        hipLaunchKernelGGL(cfl_dt_final<1>, dim3(1), dim3(kFinalBlock), 0, s->stream, s->k64.partial, nb, out, dt_dst);
    }
}

int integrals64(Solver *s, double *out_dev) {
    const int nb = nblocks(s->L.nc);
    if (s->L.D == 2)
        hipLaunchKernelGGL(integrals64<2>, dim3(nb), dim3(kBlock), 0, s->stream, s->k64);
    else
        hipLaunchKernelGGL(integrals64<3>, dim3(nb), dim3(kBlock), 0, s->stream, s->k64);
    hipLaunchKernelGGL(reduce_sum, dim3(1), dim3(kBlock), 0, s->stream, s->k64.partial, nb, s->L.D + 2, out_dev, 0);
    return 0;
}

// dispatch helpers over dimension
#define LFM_DISPATCH_D(s, call2, call3) \
    do {                                \
        if ((s)->L.D == 2) {            \
            call2;                      \
        } else {                        \
            call3;                      \
        }                               \
    } while (0)

// ---------------------------------------------------------------------------
// precision-generic stage pieces
// ---------------------------------------------------------------------------
void Solver::bc() {
    if (fp32) {
        LFM_DISPATCH_D(this, launch_bc32<2>(this), launch_bc32<3>(this));
    } else {
        LFM_DISPATCH_D(this, launch_bc64<2>(this), launch_bc64<3>(this));
    }
}

void Solver::gradvis(bool full, bool fin) {
    if (fp32) {
        LFM_DISPATCH_D(this, launch_gradvis32<2>(this, full, fin), launch_gradvis32<3>(this, full, fin));
    } else {
        LFM_DISPATCH_D(this, launch_gradvis64<2>(this, full), launch_gradvis64<3>(this, full));
    }
}

void Solver::flux() {
    if (fp32) {
        LFM_DISPATCH_D(this, launch_flux32<2>(this), launch_flux32<3>(this));
    } else {
        LFM_DISPATCH_D(this, launch_flux64<2>(this), launch_flux64<3>(this));
    }
}

void Solver::update(double A, double B, bool zero_dq, bool res, bool accumulate_res, const CflJob *cfl) {
    if (fp32) {
        LFM_DISPATCH_D(this, launch_update32<2>(this, A, B, zero_dq, res, accumulate_res, cfl),
                       launch_update32<3>(this, A, B, zero_dq, res, accumulate_res, cfl));
    } else {
        LFM_DISPATCH_D(this, launch_update64<2>(this, A, B, zero_dq, res, accumulate_res, cfl),
                       launch_update64<3>(this, A, B, zero_dq, res, accumulate_res, cfl));
    }
}

// the single-block final reduction of the block extrema an update with a
// CFL job left in the reduction scratch (shared by both precision paths)
void Solver::cfl_final(const CflJob &job) {
    const int nb = nblocks(L.nc);
    if (job.mode == 0)
        hipLaunchKernelGGL(cfl_dt_final<0>, dim3(1), dim3(kFinalBlock), 0, stream, k64.partial, nb, job.out,
                           job.dt_dst);
    else
        hipLaunchKernelGGL(cfl_dt_final<1>, dim3(1), dim3(kFinalBlock), 0, stream, k64.partial, nb, job.out,
                           job.dt_dst);
}

void Solver::cfl_dt(int mode, double param, double *out, double *dt_dst) {
    if (fp32) {
        LFM_DISPATCH_D(this, launch_cfl_dt32<2>(this, mode, param, out, dt_dst),
                       launch_cfl_dt32<3>(this, mode, param, out, dt_dst));
    } else {
        LFM_DISPATCH_D(this, launch_cfl_dt64<2>(this, mode, param, out, dt_dst),
                       launch_cfl_dt64<3>(this, mode, param, out, dt_dst));
    }
}

void Solver::prof_mark(int grp, bool start) {
    if (!prof) return;
    if (prof_n >= prof_ev.size()) return;
    (void)hipEventRecord(prof_ev[prof_n], stream);
    prof_grp[prof_n] = start ? grp : -1 - grp;
    prof_n++;
}

void Solver::stage_a(int rk) {
    (void)rk;
    // the update kernel refreshes the boundary ghosts; a full fill is needed
    // only after the state was set from outside
    if (ghosts_stale) {
        prof_mark(0, true);
        bc();
        prof_mark(0, false);
        ghosts_stale = false;
    }
    // a carried CFL reduction is taken by the step's first gradient launch
    const bool fin = cfl_pending && carry_target() == 2;
    if (fin) cfl_pending = false;
    prof_mark(1, true);
    gradvis(false, fin);
    prof_mark(1, false);
}

Solver::CflJob Solver::cfl_job(bool own_dt) {
    if (prm.adjust_dt) return CflJob{0, prm.max_co, d_dtmin, own_dt ? d_dt_ptr() : nullptr};
    return CflJob{1, dt, d_cfl, nullptr};
}

void Solver::stage_b(int rk, bool fuse_cfl, bool own_dt) {
    const bool res = prm.have_res != 0 && rk == 0;
    prof_mark(2, true);
    flux();
    prof_mark(2, false);
    // the update refreshes the boundary ghosts from the updated cells; the
    // wall forces after this step read them as this stage began
    if (post_forces && rk == nstages - 1) post_snapshot();
    // fused where measured faster: the FP64 3D update runs at 4 waves per
    // SIMD, where the appended slot gathers cost more than their own launch
    // (4.2 M-cell O-grid: 1.760 -> 1.779 ms per step; 2D and FP32 gain
    // 7-11 us per step on 1 M cells)
    if (fuse_cfl && rk == nstages - 1 && !res && (fp32 || L.D == 2)) {
        const CflJob job = cfl_job(own_dt);
        prof_mark(5, true);
        update(A[rk], B[rk], rk == 0, false, false, &job);
        prof_mark(5, false);
        final_or_carry(job);
        cfl_fused = true;
    } else {
        prof_mark(3, true);
        update(A[rk], B[rk], rk == 0, res, false);
        prof_mark(3, false);
    }
}

void Solver::end_step(bool own_dt) {
    if (cfl_fused) {
        cfl_fused = false;
    } else {
        prof_mark(4, true);
        const CflJob job = cfl_job(own_dt);
        cfl_dt(job.mode, job.param, job.out, job.dt_dst);
        prof_mark(4, false);
    }
    steps_done++;
    if (prm.have_res) res_valid = true;
}

// the per-cell face slots (each cell's first GatherUnroll gather entries:
// other cell, area vector, weight), read by the one-launch stages and the
// FP32 gradient kernel; allocated on first use (under the set-up mutex, as a
// capture on another thread requires) and filled from the gather list on
// the handle's stream.  When memory cannot be had the kernels read the
// gather list and the face arrays, which gives the same bits.
bool Solver::ensure_slots32() {
    if (!fp32 || slots_failed) return false;
    if (k32->sl_nb) return true;
    const int U = L.D == 2 ? GatherUnroll<2>::value : GatherUnroll<3>::value;
    std::lock_guard<std::mutex> lock(device_setup_mutex());
    void *p[3] = {nullptr, nullptr, nullptr};
    const size_t bytes[3] = {slot_array_len(L.nc, U, 1) * sizeof(int32_t),
                             slot_array_len(L.nc, U, L.D) * sizeof(float), slot_array_len(L.nc, U, 1) * sizeof(float)};
    for (int i = 0; i < 3; i++)
        if (hipMalloc(&p[i], bytes[i]) != hipSuccess) {
            (void)hipGetLastError();
            for (int k = 0; k < i; k++) (void)hipFree(p[k]);
            slots_failed = true;
            return false;
        }
    for (int i = 0; i < 3; i++) {
        allocs.push_back(p[i]);
        device_bytes += (int64_t)bytes[i];
    }
    LFM_DISPATCH_D(this, launch_build_slots32<2>(this, (int32_t *)p[0], (float *)p[1], (float *)p[2]),
                   launch_build_slots32<3>(this, (int32_t *)p[0], (float *)p[1], (float *)p[2]));
    k32->sl_nb = (const int32_t *)p[0];
    k32->sl_S = (const float *)p[1];
    k32->sl_w = (const float *)p[2];
    return true;
}

// the second state buffer of the one-launch stages (and the slots they
// read), allocated on first use; without them the step runs the separate
// launches, which compute the same bits
bool Solver::ensure_alt32() {
    if (!fused_ok) return false;
    if (q_alt32) return true;
    if (!ensure_slots32()) {
        fused_ok = false;
        return false;
    }
    const size_t nq = (size_t)(L.D + 2) * ((size_t)L.nc + L.ng);
    std::lock_guard<std::mutex> lock(device_setup_mutex());
    void *p = nullptr;
    if (hipMalloc(&p, nq * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        fused_ok = false;
        return false;
    }
    allocs.push_back(p);
    device_bytes += (int64_t)(nq * sizeof(float));
    q_alt32 = (float *)p;
    return true;
}

// the first stage run as one fused launch (nstages when none is): every
// stage, except a residual stage 0, which keeps the separate launches for
// its per-block sums
int Solver::fused_from() const {
    if (!fused_ok || post_forces) return nstages;
    const int first = prm.have_res ? 1 : 0;
    return first < nstages ? first : nstages;
}

void Solver::stage_fused(int rk, bool from_alt) {
    stage_a(rk);  // the ghost fill after an outside write (no gradients here)
    float *qin = from_alt ? q_alt32 : k32->q, *qout = from_alt ? k32->q : q_alt32;
    if (rk == nstages - 1) {
        const CflJob job = cfl_job(true);
        prof_mark(5, true);
        LFM_DISPATCH_D(this, launch_stage_fused32<2>(this, A[rk], B[rk], rk == 0, qin, qout, &job, false),
                       launch_stage_fused32<3>(this, A[rk], B[rk], rk == 0, qin, qout, &job, false));
        prof_mark(5, false);
        final_or_carry(job);
        cfl_fused = true;
    } else {
        // a carried reduction is taken by the step's first stage
        const bool fin = cfl_pending && carry_target() == 1;
        if (fin) cfl_pending = false;
        prof_mark(3, true);
        LFM_DISPATCH_D(this, launch_stage_fused32<2>(this, A[rk], B[rk], rk == 0, qin, qout, nullptr, fin),
                       launch_stage_fused32<3>(this, A[rk], B[rk], rk == 0, qin, qout, nullptr, fin));
        prof_mark(3, false);
    }
}

// the launch of a step that can take the previous step's carried CFL
// reduction before anything writes the reduction scratch: 1, the first
// one-launch stage (not a residual stage, and not the only stage, which
// forms the CFL itself); 2, the first FP32 gradient launch (each stage of a
// viscous or M2AUSM solver has one, before its update); 0, none
int Solver::carry_target() const {
    if (!L.halo.empty()) return 0;
    if (fused_ok) return q_alt32 && fused_from() == 0 && nstages > 1 ? 1 : 0;
    return fp32 && (visc || prm.solver == 2) ? 2 : 0;
}

// the final reduction of a CFL job the last stage formed: carried into the
// next step (compute_cfl inside a run that steps again) or launched now
void Solver::final_or_carry(const CflJob &job) {
    if (defer_cfl && job.mode == 1 && job.out == d_cfl && carry_target() != 0) {
        cfl_pending = true;
        return;
    }
    prof_mark(4, true);
    cfl_final(job);
    prof_mark(4, false);
}

void Solver::flush_cfl() {
    if (!cfl_pending) return;
    cfl_pending = false;
    prof_mark(4, true);
    cfl_final(CflJob{1, dt, d_cfl, nullptr});
    prof_mark(4, false);
}

// one full time step (Mesh::solve loop body), no host synchronisation
void Solver::step() {
    const int first = fused_ok && ensure_alt32() ? fused_from() : nstages;
    if (carry_target() == 0) flush_cfl();
    for (int rk = 0; rk < first; rk++) {
        stage_a(rk);
        stage_b(rk, true, true);
    }
    bool in_alt = false;
    for (int rk = first; rk < nstages; rk++) {
        stage_fused(rk, in_alt);
        in_alt = !in_alt;
    }
    // an odd number of one-launch stages left the state in the second
    // buffer: it becomes the state buffer (kernels take k32 by value at
    // launch, so later launches see the swap)
    if (in_alt) std::swap(k32->q, q_alt32);
    end_step(true);
}

}  // namespace lfm

using lfm::fail;
using lfm::Solver;


extern "C" {

const char *lfm_last_error(void) { return lfm::g_last_error.c_str(); }

int lfm_solver_create(const lfm_polymesh *mesh, const lfm_params *prm, const double *p, const double *T,
                      const double *U3, const double *alpha, lfm_solver **out) {
    return lfm_solver_create_part(mesh, nullptr, prm, p, T, U3, alpha, out);
}

int lfm_solver_create_part(const lfm_polymesh *mesh, const lfm_partition *part, const lfm_params *prm,
                           const double *p, const double *T, const double *U3, const double *alpha,
                           lfm_solver **out) {
    if (!mesh || !prm || !p || !T || !U3 || !out) return fail(LFM_EINVAL, "null argument");
    *out = nullptr;
    if (prm->solver < 0 || prm->solver > 2) return fail(LFM_EINVAL, "solver must be 0 (M1), 1 (M2) or 2 (M2AUSM)");
    if (prm->rk_order < 0 || prm->rk_order > 5) return fail(LFM_EINVAL, "rkOrder must be 0..5");
    if (prm->precision != LFM_PREC_FP64 && prm->precision != LFM_PREC_FP32P)
        return fail(LFM_EINVAL, "precision must be 64 or 32");
    auto s = std::make_unique<Solver>();
    s->prm = *prm;
    std::string err;
    try {
        err = lfm::build_layout(*mesh, part, *prm, p, T, U3, alpha, s->L);
    } catch (const std::bad_alloc &) {
        return fail(LFM_ENOMEM, "out of host memory building the solver layout");
    } catch (const std::exception &e) {
        return fail(LFM_EINVAL, std::string("building the solver layout: ") + e.what());
    }
    if (!err.empty()) return fail(LFM_EINVAL, err);
    s->gas = lfm::derive_gas(*prm);
    s->nstages = lfm::rk_tables(prm->rk_order, s->A, s->B);
    s->dt = prm->dt;
    s->fp32 = prm->precision == LFM_PREC_FP32P;
    // the Smagorinsky stress does not scale with mu
    s->visc = prm->mu != 0.0 || !prm->laminar;
    std::lock_guard<std::mutex> lock(lfm::device_setup_mutex());
    HIP_TRY(hipGetDevice(&s->device));
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&s->ev0));
    HIP_TRY(hipEventCreate(&s->ev1));
    HIP_TRY(hipHostMalloc((void **)&s->h_scalar, 64 * sizeof(double), hipHostMallocDefault));
    int rc = s->fp32 ? lfm::create_fp32p(s.get()) : lfm::create_fp64(s.get());
    if (rc) return rc;
    s->halo_dev.resize(s->L.halo.size());
    for (size_t i = 0; i < s->L.halo.size(); i++) {
        const auto &hn = s->L.halo[i];
        auto &hd = s->halo_dev[i];
        hd.n_send = (int)hn.send_cells.size();
        hd.n_recv = (int)hn.ghost.size();
        std::vector<int32_t> gslot(hn.ghost.size());
        for (size_t j = 0; j < hn.ghost.size(); j++) gslot[j] = s->L.nc + hn.ghost[j];
        if (lfm::dev_upload(s.get(), &hd.send_cells, hn.send_cells.data(), hn.send_cells.size()) ||
            lfm::dev_upload(s.get(), &hd.recv_ghost, gslot.data(), gslot.size()) ||
            lfm::dev_zeros(s.get(), &hd.recv_src, gslot.size()))
            return LFM_EDEVICE;
    }
    HIP_TRY(hipDeviceSynchronize());
    auto *h = new lfm_solver;
    h->impl = std::move(s);
    *out = h;
    return LFM_OK;
}

int lfm_solver_destroy(lfm_solver *h) {
    if (!h) return LFM_OK;
    if (h->impl && h->impl->stream) (void)hipStreamSynchronize(h->impl->stream);
    delete h;
    return LFM_OK;
}

int lfm_get_info(lfm_solver *h, lfm_info *info) {
    if (!h || !info) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    HIP_TRY(hipStreamSynchronize(s->stream));
    info->dim = s->L.D;
    info->n_cells = s->L.nc;
    info->n_cells_sm0 = s->L.nsm0;
    info->n_faces = s->L.nf;
    info->n_ghosts = s->L.ng;
    info->max_faces_per_cell = s->L.max_faces;
    info->precision = s->prm.precision;
    info->solver = s->prm.solver;
    info->device_bytes = s->device_bytes;
    HIP_TRY(hipMemcpy(s->h_scalar, s->d_scalar, 16 * sizeof(double), hipMemcpyDeviceToHost));
    info->dt = s->h_scalar[0];
    info->cfl = s->h_scalar[8];
    info->steps_done = s->steps_done;
    return LFM_OK;
}

int lfm_prepare_for_timestep(lfm_solver *h) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    Solver *s = h->impl.get();
    s->zero_state_for_step();
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_prepare_for_rkstep(lfm_solver *h, int rk) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    Solver *s = h->impl.get();
    if (rk < 0 || rk >= std::max(1, s->nstages) || rk > 4) return fail(LFM_EINVAL, "rk stage out of range");
    s->pending_A = s->A[rk];
    s->pending_rk = rk;
    return LFM_OK;
}

int lfm_set_boundary_conditions(lfm_solver *h) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    h->impl->bc();
    h->impl->ghosts_stale = false;
    if (h->impl->post_forces) h->impl->post_snapshot();
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_calc_gradients(lfm_solver *h) {
    // Gradients are produced together with the viscous terms by the same
    // cell gather (lfm_calc_vis); calling this alone is a no-op.
    if (!h) return fail(LFM_EINVAL, "null handle");
    return LFM_OK;
}

int lfm_calc_vis(lfm_solver *h) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    if (!h->impl->halo_ready()) return fail(LFM_ESTATE, "halo set-up not imported for every neighbour");
    h->impl->gradvis(true);
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_one_rk_step(lfm_solver *h, int rk, double dt, double *res) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    Solver *s = h->impl.get();
    if (rk < 0 || rk > 4) return fail(LFM_EINVAL, "rk stage out of range");
    if (!s->halo_ready()) return fail(LFM_ESTATE, "halo set-up not imported for every neighbour");
    const double B = s->B[rk];
    s->h_scalar[0] = dt;
    HIP_TRY(hipMemcpyAsync(s->d_dt_ptr(), s->h_scalar, sizeof(double), hipMemcpyHostToDevice, s->stream));
    s->dt = dt;
    const bool want_res = res && rk == 0 && s->prm.have_res;
    s->flux();
    s->update(s->pending_A, B, false, want_res, false);
    s->pending_A = 1.0;
    HIP_TRY(hipGetLastError());
    if (want_res) {
        HIP_TRY(hipMemcpyAsync(s->h_scalar + 16, s->d_res, (s->L.D + 2) * sizeof(double), hipMemcpyDeviceToHost,
                               s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (int k = 0; k < s->L.D + 2; k++) res[k] += s->h_scalar[16 + k];
        s->res_valid = true;
    }
    return LFM_OK;
}

int lfm_compute_dt(lfm_solver *h, double max_co, double *dt_out) {
    if (!h || !dt_out) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    s->cfl_dt(0, max_co, s->d_dtmin, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(s->h_scalar + 9, s->d_dtmin, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *dt_out = s->h_scalar[9];
    return LFM_OK;
}

int lfm_compute_cfl(lfm_solver *h, double dt, double *cfl_out) {
    if (!h || !cfl_out) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    s->cfl_dt(1, dt, s->d_cfl, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(s->h_scalar + 8, s->d_cfl, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *cfl_out = s->h_scalar[8];
    return LFM_OK;
}

int lfm_run(lfm_solver *h, int nsteps) {
    if (!h || nsteps < 0) return fail(LFM_EINVAL, "bad argument");
    Solver *s = h->impl.get();
    if (!s->L.halo.empty())
        return fail(LFM_ESTATE, "a partition with processor patches steps through lfm_part_group_run or the "
                                "per-stage calls with halo exchange");
    for (int i = 0; i < nsteps; i++) {
        s->defer_cfl = i + 1 < nsteps && !s->post;
        s->step();
        if (s->post) {
            const int rc = s->post_step();
            if (rc) return rc;
        }
    }
    s->flush_cfl();
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_synchronize(lfm_solver *h) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    HIP_TRY(hipStreamSynchronize(h->impl->stream));
    return LFM_OK;
}

int lfm_time_steps(lfm_solver *h, int nsteps, double *ms_out) {
    if (!h || !ms_out || nsteps < 0) return fail(LFM_EINVAL, "bad argument");
    Solver *s = h->impl.get();
    if (!s->L.halo.empty())
        return fail(LFM_ESTATE, "a partition with processor patches steps through lfm_part_group_run or the "
                                "per-stage calls with halo exchange");
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    for (int i = 0; i < nsteps; i++) {
        s->defer_cfl = i + 1 < nsteps;
        s->step();
    }
    s->flush_cfl();
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    HIP_TRY(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *ms_out = ms;
    return LFM_OK;
}

int lfm_profile_steps(lfm_solver *h, int nsteps, double *ms_per_group, double *ms_total) {
    if (!h || !ms_per_group || !ms_total || nsteps < 1) return fail(LFM_EINVAL, "bad argument");
    Solver *s = h->impl.get();
    if (!s->L.halo.empty())
        return fail(LFM_ESTATE, "a partition with processor patches steps through lfm_part_group_run or the "
                                "per-stage calls with halo exchange");
    const size_t need = (size_t)nsteps * (s->nstages * 8 + 2);
    for (hipEvent_t e : s->prof_ev) (void)hipEventDestroy(e);
    s->prof_ev.assign(need, nullptr);
    s->prof_grp.assign(need, 0);
    for (size_t i = 0; i < need; i++) HIP_TRY(hipEventCreate(&s->prof_ev[i]));
    s->prof_n = 0;
    s->prof = true;
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    for (int i = 0; i < nsteps; i++) {
        s->defer_cfl = i + 1 < nsteps;
        s->step();
    }
    s->flush_cfl();
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    s->prof = false;
    HIP_TRY(hipEventSynchronize(s->ev1));
    for (int g = 0; g < Solver::kProfGroups; g++) ms_per_group[g] = 0.0;
    for (size_t i = 0; i + 1 < s->prof_n; i += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->prof_ev[i], s->prof_ev[i + 1]));
        const int g = s->prof_grp[i];
        if (g >= 0 && g < Solver::kProfGroups) ms_per_group[g] += ms;
    }
    float tot = 0.f;
    HIP_TRY(hipEventElapsedTime(&tot, s->ev0, s->ev1));
    *ms_total = tot;
    for (hipEvent_t e : s->prof_ev) (void)hipEventDestroy(e);
    s->prof_ev.clear();
    s->prof_grp.clear();
    return LFM_OK;
}

int lfm_get_field(lfm_solver *h, int which, double *out) {
    if (!h || !out) return fail(LFM_EINVAL, "null argument");
    return h->impl->get_field(which, out);
}

int lfm_set_field(lfm_solver *h, int which, const double *in) {
    if (!h || !in) return fail(LFM_EINVAL, "null argument");
    return h->impl->set_field(which, in);
}

int lfm_get_residual(lfm_solver *h, double *res) {
    if (!h || !res) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    HIP_TRY(hipMemcpyAsync(s->h_scalar + 16, s->d_res, (s->L.D + 2) * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < s->L.D + 2; k++) res[k] = s->h_scalar[16 + k];
    return LFM_OK;
}

int lfm_conservation(lfm_solver *h, double *out) {
    if (!h || !out) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    lfm::integrals(s, s->d_scalar + 10);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(s->h_scalar + 24, s->d_scalar + 10, (s->L.D + 2) * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int k = 0; k < s->L.D + 2; k++) out[k] = s->h_scalar[24 + k];
    return LFM_OK;
}

int lfm_check_finite(lfm_solver *h, int64_t *n_bad) {
    if (!h || !n_bad) return fail(LFM_EINVAL, "null argument");
    return h->impl->check_finite(n_bad);
}

int lfm_check_finite_noalloc(lfm_solver *h, int64_t *n_bad) {
    if (!h || !n_bad) return fail(LFM_EINVAL, "null argument");
    return h->impl->check_finite_noalloc(n_bad);
}

}  // extern "C"
