// solver_fp32p.hip -- creation and launches of the FP32-perturbation path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "kernels_fp32p.hpp"
#include "solver_state.hpp"

namespace lfm {

static constexpr int kBlock = 256;
static inline int nblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
static int up(Solver *s, T **dst, const T *src, size_t n) {
    if (n == 0) {
        *dst = nullptr;
        return 0;
    }
    HIP_TRY(hipMalloc((void **)dst, n * sizeof(T)));
    s->allocs.push_back((void *)*dst);
    s->device_bytes += (int64_t)(n * sizeof(T));
    if (src)
        HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    else
        HIP_TRY(hipMemset(*dst, 0, n * sizeof(T)));
    return 0;
}

static std::vector<float> to_f32(const std::vector<double> &v) {
    std::vector<float> o(v.size());
    for (size_t i = 0; i < v.size(); i++) o[i] = (float)v[i];
    return o;
}

int create_fp32p(Solver *s) {
    const SolverLayout &L = s->L;
    const int D = L.D;
    const size_t nc = L.nc, ntot = (size_t)L.nc + L.ng, nf = L.nf;
    s->k32 = new KArgs32();
    KArgs32 &a = *s->k32;
    std::memset(&a, 0, sizeof(a));
    a.nc = L.nc;
    a.nsm0 = L.nsm0;
    a.ng = L.ng;
    a.nf = L.nf;
    a.ng_bnd = L.ng_bnd;
    a.ntot = (int)ntot;
    const Gas &g = s->gas;
    // freestream constants (FP64 derivations, rounded once)
    double m_inf[3] = {g.rho_inf * g.U_inf[0], g.rho_inf * g.U_inf[1], g.rho_inf * g.U_inf[2]};
    const double rhoE_inf = g.rho_inf * g.E_inf;
    double uinf2 = 0.0;
    for (int k = 0; k < D; k++) uinf2 += g.U_inf[k] * g.U_inf[k];
    const double Rpsi_inf = g.p_inf / g.rho_inf;
    const double gog = g.gamma / g.gm1;
    const double H_inf = gog * Rpsi_inf + 0.5 * uinf2;
    a.gamma = (float)g.gamma;
    a.gm1 = (float)g.gm1;
    a.gog = (float)gog;
    a.R_inv = (float)g.R_inv;
    a.mu = (float)g.mu;
    a.cp = (float)g.cp;
    a.pr_inv = (float)g.pr_inv;
    a.rho_inf = (float)g.rho_inf;
    a.p_inf = (float)g.p_inf;
    a.T_inf = (float)g.T_inf;
    a.Rpsi_inf = (float)Rpsi_inf;
    a.E_inf = (float)g.E_inf;
    a.H_inf = (float)H_inf;
    a.G_inf = (float)(g.rho_inf * H_inf);
    a.c_inf = (float)std::sqrt(g.gamma * Rpsi_inf);
    for (int k = 0; k < 3; k++) {
        a.uinf[k] = k < D ? (float)g.U_inf[k] : 0.f;
        a.minf[k] = k < D ? (float)m_inf[k] : 0.f;
        a.m_inf_d[k] = k < D ? m_inf[k] : 0.0;
    }
    a.rho_inf_d = g.rho_inf;
    a.rhoE_inf_d = rhoE_inf;

    int rc = 0;
    rc |= up(s, (int32_t **)&a.fP, L.fP.data(), nf);
    rc |= up(s, (int32_t **)&a.fN, L.fN.data(), nf);
    {
        // S as one record per face ([nf][D]); d and w per component
        std::vector<float> S((size_t)D * nf);
        for (size_t f = 0; f < nf; f++)
            for (int k = 0; k < D; k++) S[f * D + k] = (float)L.S[(size_t)k * nf + f];
        auto d = to_f32(L.d), w = to_f32(L.w);
        rc |= up(s, (float **)&a.S, S.data(), S.size());
        rc |= up(s, (float **)&a.d, d.data(), d.size());
        rc |= up(s, (float **)&a.w, w.data(), w.size());
        auto vinv = to_f32(L.vinv), sig = to_f32(L.sigma);
        rc |= up(s, (float **)&a.vinv, vinv.data(), nc);
        bool sponge = false;
        for (float v : sig) sponge |= v != 0.f;
        if (sponge) rc |= up(s, (float **)&a.sigma, sig.data(), nc);
        if (!L.smag.empty()) {
            auto sm = to_f32(L.smag);
            rc |= up(s, (float **)&a.smag, sm.data(), nc);
        }
    }
    rc |= up(s, (int32_t **)&a.g_off, L.g_off.data(), nc + 1);
    rc |= up(s, (int32_t **)&a.g_ent, L.g_ent.data(), L.g_ent.size());
    rc |= up(s, (int32_t **)&a.ghost_cell, L.ghost_cell.data(), L.ng);
    rc |= up(s, (int32_t **)&a.ghost_role, L.ghost_role.data(), L.ng);
    rc |= up(s, (double **)&a.vol_d, L.volume.data(), nc);
    // initial deviations; unset-patch ghosts hold q = 0 (q' = -q_inf)
    {
        std::vector<float> q0((size_t)(D + 2) * ntot, 0.f);
        const double qinf[5] = {g.rho_inf, m_inf[0], m_inf[1], m_inf[2], rhoE_inf};
        double qi[5];
        qi[0] = qinf[0];
        for (int k = 0; k < D; k++) qi[k + 1] = qinf[k + 1];
        qi[D + 1] = rhoE_inf;
        for (int k = 0; k < D + 2; k++) {
            for (size_t c = 0; c < nc; c++) q0[q32_index(k, c, ntot)] = (float)(L.q0[k * nc + c] - qi[k]);
            for (size_t gi = 0; gi < (size_t)L.ng; gi++) q0[q32_index(k, nc + gi, ntot)] = (float)(-qi[k]);
        }
        rc |= up(s, &a.q, q0.data(), q0.size());
        // one-launch stages (a second state buffer, allocated on the first
        // step that uses it; its unset-patch ghosts are copied from q by the
        // first fused stage)
        s->fused_ok = !s->visc && s->prm.solver < 2 && L.halo.empty();
    }
    rc |= up(s, &a.dq, (float *)nullptr, (size_t)(D + 2) * nc);
    rc |= up(s, &a.res_cell, (float *)nullptr, (size_t)(D + 2) * nc);
    // gradient records (GradRec32): dudx, dTdx and (M2AUSM) rho / p gradients
    const bool rg = s->prm.solver == 2;
    a.gstride = grad_record_stride(D, rg);
    rc |= up(s, &a.dudx, (float *)nullptr, (size_t)a.gstride * ntot);
    // tauMC / sigmaU are kept only for the halo payload of a partition (the
    // face kernel otherwise recomputes them from dudx); U_grad shares dudx's
    // storage (equal bit for bit) except where halo ghosts must read zero
    // U_grad, or where no gradient pass writes it (solver 2 without minmod)
    const bool part = !L.halo.empty();
    if (part) {
        rc |= up(s, &a.tauMC, (float *)nullptr, (size_t)D * D * ntot);
        rc |= up(s, &a.sigmaU, (float *)nullptr, (size_t)D * ntot);
    }
    if (rc) return LFM_EDEVICE;
    a.dTdx = a.dudx + D * D;
    if (s->prm.solver == 2) {
        a.rho_grad = a.dudx + D * D + D;
        a.p_grad = a.dudx + D * D + 2 * D;
        if (part || !s->prm.minmod)
            rc |= up(s, &a.U_grad, (float *)nullptr, (size_t)D * D * ntot);
        else
            a.U_grad = a.dudx;
    }
    rc |= up(s, &a.frhs, (float *)nullptr, (size_t)(D + 2) * nf);
    const size_t npart = (size_t)nblocks((long)std::max(nc, nf)) * (D + 2) + 16;
    rc |= up(s, &a.partial, (double *)nullptr, npart);
    rc |= up(s, &s->d_scalar, (double *)nullptr, 16);
    if (rc) return LFM_EDEVICE;
    a.dt = s->d_scalar;
    s->d_res = s->d_scalar + 2;
    s->d_cfl = s->d_scalar + 8;
    s->d_dtmin = s->d_scalar + 9;
    s->k64.partial = a.partial;  // shared reduction scratch for the final kernels
    a.d_cfl = s->d_cfl;
    a.d_dtmin = s->d_dtmin;
    a.nblk_cells = nblocks((long)nc);
    HIP_TRY(hipMemcpy(s->d_scalar, &s->dt, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

template <int D>
void launch_bc32(Solver *s) {
    if (s->L.ng) hipLaunchKernelGGL(bc_fill32<D>, dim3(nblocks(s->L.ng)), dim3(kBlock), 0, s->stream, *s->k32);
}

template <int D, bool V, bool SL>
static void launch_gradvis32_v(Solver *s, bool fin) {
    const int nb = nblocks(s->L.nc);
    const dim3 grid(nb + (fin ? 1 : 0));
    const int fb = fin ? nb : 0;
    if (s->prm.solver == 2 && s->prm.minmod)
        hipLaunchKernelGGL((grad_vis32<D, 2, V, true, SL>), grid, dim3(kBlock), 0, s->stream, *s->k32, fb);
    else if (s->prm.solver == 2)
        hipLaunchKernelGGL((grad_vis32<D, 0, V, true, SL>), grid, dim3(kBlock), 0, s->stream, *s->k32, fb);
    else
        hipLaunchKernelGGL((grad_vis32<D, 0, V, false, SL>), grid, dim3(kBlock), 0, s->stream, *s->k32, fb);
}

// 3D gradients read the face slots (same bits as the gather list; 4.2 M-cell
// O-grid 0.98 -> 0.86 ms per step).  2D gradients keep the list: there the
// slot streams cost more than the cached face records they replace (1 M-cell
// annulus 0.114 -> 0.134 ms per step with slots).  fin: the launch also
// reduces the previous step's CFL maxima in one extra block (the caller
// knows the launch happens: full, viscous or M2AUSM).
template <int D>
void launch_gradvis32(Solver *s, bool full, bool fin) {
    if (!full && !s->visc && s->prm.solver != 2) return;
    const bool V = full || s->visc;
    if constexpr (D == 3) {
        if (s->ensure_slots32()) {
            V ? launch_gradvis32_v<D, true, true>(s, fin) : launch_gradvis32_v<D, false, true>(s, fin);
            return;
        }
    }
    V ? launch_gradvis32_v<D, true, false>(s, fin) : launch_gradvis32_v<D, false, false>(s, fin);
}

template <int D, bool V, bool ST>
static void launch_flux32_v(Solver *s) {
    const int nb = nblocks(s->L.nf);
    switch (s->prm.solver) {
    case 0: hipLaunchKernelGGL((face_flux32<D, 0, V, ST>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32); break;
    case 1: hipLaunchKernelGGL((face_flux32<D, 1, V, ST>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32); break;
    default: hipLaunchKernelGGL((face_flux32<D, 2, V, ST>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32); break;
    }
}

template <int D>
void launch_flux32(Solver *s) {
    if (s->visc)
        s->k32->tauMC ? launch_flux32_v<D, true, true>(s) : launch_flux32_v<D, true, false>(s);
    else
        launch_flux32_v<D, false, false>(s);
}

template <int D>
void launch_update32(Solver *s, double A, double B, bool zero_dq, bool res, bool accumulate_res,
                     const Solver::CflJob *cfl) {
    const int nb = nblocks(s->L.nc);
    if (res) {
        hipLaunchKernelGGL((cell_update32<D, true, -1>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32, (float)A,
                           (float)B, (int)zero_dq, 0.0);
        hipLaunchKernelGGL(reduce_sum, dim3(1), dim3(kBlock), 0, s->stream, s->k32->partial, nb, D + 2, s->d_res,
                           (int)accumulate_res);
    } else if (cfl && cfl->mode == 0) {
        hipLaunchKernelGGL((cell_update32<D, false, 0>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32, (float)A,
                           (float)B, (int)zero_dq, cfl->param);
    } else if (cfl) {
        hipLaunchKernelGGL((cell_update32<D, false, 1>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32, (float)A,
                           (float)B, (int)zero_dq, cfl->param);
    } else {
        hipLaunchKernelGGL((cell_update32<D, false, -1>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32, (float)A,
                           (float)B, (int)zero_dq, 0.0);
    }
}

// fin: one block more reduces the CFL maxima of the previous step's last
// stage (left in a.partial) into a.d_cfl
template <int D, int SOLVER>
static void launch_fused32_v(Solver *s, float A, float B, int zero_dq, const KArgs32 &a, float *qout,
                             const Solver::CflJob *cfl, bool fin) {
    const int nb = nblocks(s->L.nc);
    if (!cfl) {
        hipLaunchKernelGGL((stage_fused32<D, SOLVER, -1>), dim3(nb + (fin ? 1 : 0)), dim3(kBlock), 0, s->stream, a,
                           qout, A, B, zero_dq, 0.0, fin ? nb : 0);
        return;
    }
    if (cfl->mode == 0)
        hipLaunchKernelGGL((stage_fused32<D, SOLVER, 0>), dim3(nb), dim3(kBlock), 0, s->stream, a, qout, A, B,
                           zero_dq, cfl->param, 0);
    else
        hipLaunchKernelGGL((stage_fused32<D, SOLVER, 1>), dim3(nb), dim3(kBlock), 0, s->stream, a, qout, A, B,
                           zero_dq, cfl->param, 0);
}

template <int D>
void launch_stage_fused32(Solver *s, double A, double B, bool zero_dq, const float *qin, float *qout,
                          const Solver::CflJob *cfl, bool fin) {
    KArgs32 a = *s->k32;
    a.q = const_cast<float *>(qin);
    if (s->prm.solver == 0)
        launch_fused32_v<D, 0>(s, (float)A, (float)B, (int)zero_dq, a, qout, cfl, fin);
    else
        launch_fused32_v<D, 1>(s, (float)A, (float)B, (int)zero_dq, a, qout, cfl, fin);
}

template <int D>
void launch_build_slots32(Solver *s, int32_t *nb, float *S, float *w) {
    hipLaunchKernelGGL(build_slots32<D>, dim3(nblocks(s->L.nc)), dim3(kBlock), 0, s->stream, *s->k32, nb, S, w);
}

template <int D>
void launch_cfl_dt32(Solver *s, int mode, double param, double *out, double *dt_dst) {
    const int nb = nblocks(s->L.nc);
    if (mode == 0) {
        hipLaunchKernelGGL((cfl_dt_cells32<D, 0>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32, param);
        hipLaunchKernelGGL(cfl_dt_final<0>, dim3(1), dim3(kFinalBlock), 0, s->stream, s->k32->partial, nb, out, dt_dst);
    } else {
        hipLaunchKernelGGL((cfl_dt_cells32<D, 1>), dim3(nb), dim3(kBlock), 0, s->stream, *s->k32, param);
        hipLaunchKernelGGL(cfl_dt_final<1>, dim3(1), dim3(kFinalBlock), 0, s->stream, s->k32->partial, nb, out, dt_dst);
    }
}

template <int D>
void launch_integrals32(Solver *s, double *out) {
    const int nb = nblocks(s->L.nc);
    hipLaunchKernelGGL(integrals32<D>, dim3(nb), dim3(kBlock), 0, s->stream, *s->k32);
    hipLaunchKernelGGL(reduce_sum, dim3(1), dim3(kBlock), 0, s->stream, s->k32->partial, nb, D + 2, out, 0);
}

template <int D>
void launch_nonfinite32(Solver *s, unsigned long long *d) {
    hipLaunchKernelGGL(count_nonfinite32<D>, dim3(nblocks(s->L.nc)), dim3(kBlock), 0, s->stream, s->k32->q, s->L.nc,
                       s->L.nc + s->L.ng, d);
}

// FP32 field access (converted to FP64 values in original cell order)
int get_field32(Solver *s, int which, double *out) {
    const SolverLayout &L = s->L;
    const int D = L.D;
    const size_t nc = L.nc, nt = (size_t)L.nc + L.ng;
    KArgs32 &a = *s->k32;
    const float *ptr = nullptr;
    int ncomp = 0;
    size_t stride = nt;
    switch (which) {
    case LFM_FIELD_Q: ptr = a.q; ncomp = D + 2; break;
    case LFM_FIELD_DQ: ptr = a.dq; ncomp = D + 2; stride = nc; break;
    case LFM_FIELD_DUDX: ptr = a.dudx; ncomp = D * D; break;
    case LFM_FIELD_DTDX: ptr = a.dTdx; ncomp = D; break;
    case LFM_FIELD_TAUMC: ptr = a.tauMC; ncomp = D * D; break;
    case LFM_FIELD_SIGMAU: ptr = a.sigmaU; ncomp = D; break;
    case LFM_FIELD_RHOGRAD: ptr = a.rho_grad; ncomp = D; break;
    case LFM_FIELD_PGRAD: ptr = a.p_grad; ncomp = D; break;
    case LFM_FIELD_UGRAD: ptr = a.U_grad; ncomp = D * D; break;
    case LFM_FIELD_RES: ptr = a.res_cell; ncomp = D + 2; stride = nc; break;
    default: return fail(LFM_EINVAL, "unknown field");
    }
    std::vector<float> derived;
    if (!ptr && (which == LFM_FIELD_TAUMC || which == LFM_FIELD_SIGMAU)) {
        // not stored: formed from dudx and the state
        float *dev = nullptr;
        HIP_TRY(hipMalloc((void **)&dev, (size_t)(D * D + D) * nc * sizeof(float)));
        if (D == 2)
            hipLaunchKernelGGL(vis_tensor_field32<2>, dim3(nblocks((long)nc)), dim3(kBlock), 0, s->stream, a, dev,
                               dev + (size_t)D * D * nc);
        else
            hipLaunchKernelGGL(vis_tensor_field32<3>, dim3(nblocks((long)nc)), dim3(kBlock), 0, s->stream, a, dev,
                               dev + (size_t)D * D * nc);
        derived.resize((size_t)(D * D + D) * nc);
        hipError_t e = hipMemcpyAsync(derived.data(), dev, derived.size() * sizeof(float), hipMemcpyDeviceToHost,
                                      s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        (void)hipFree(dev);
        HIP_TRY(e);
        const float *src = derived.data() + (which == LFM_FIELD_TAUMC ? 0 : (size_t)D * D * nc);
        for (int k = 0; k < ncomp; k++)
            for (size_t ci = 0; ci < nc; ci++) out[k * nc + L.orig[ci]] = (double)src[k * nc + ci];
        return LFM_OK;
    }
    if (!ptr) return fail(LFM_EINVAL, "field not allocated for this solver");
    const bool in_record = ptr != a.q && ptr != a.dq && ptr != a.res_cell && ptr != a.tauMC && ptr != a.sigmaU &&
                           (ptr != a.U_grad || a.U_grad == a.dudx);
    if (in_record) {
        // component k of cell ci at ptr[ci * gstride + k]
        std::vector<float> recs((size_t)a.gstride * nc);
        HIP_TRY(hipMemcpy(recs.data(), a.dudx, recs.size() * sizeof(float), hipMemcpyDeviceToHost));
        const size_t off = (size_t)(ptr - a.dudx);
        for (int k = 0; k < ncomp; k++)
            for (size_t ci = 0; ci < nc; ci++)
                out[k * nc + L.orig[ci]] = (double)recs[ci * a.gstride + off + k];
        return LFM_OK;
    }
    double qinf[5] = {a.rho_inf_d, a.m_inf_d[0], a.m_inf_d[1], a.m_inf_d[2], 0.0};
    if (which == LFM_FIELD_Q) {
        // cell records: components 0-3 interleaved, component 4 after them
        qinf[D + 1] = a.rhoE_inf_d;
        std::vector<float> rec((size_t)4 * nc), e4(D == 3 ? nc : 0);
        HIP_TRY(hipMemcpy(rec.data(), a.q, rec.size() * sizeof(float), hipMemcpyDeviceToHost));
        if (D == 3) HIP_TRY(hipMemcpy(e4.data(), a.q + 4 * nt, nc * sizeof(float), hipMemcpyDeviceToHost));
        for (int k = 0; k < D + 2; k++)
            for (size_t ci = 0; ci < nc; ci++)
                out[k * nc + L.orig[ci]] = qinf[k] + (double)(k < 4 ? rec[4 * ci + k] : e4[ci]);
        return LFM_OK;
    }
    qinf[D + 1] = a.rhoE_inf_d;
    std::vector<float> tmp(nc);
    for (int k = 0; k < ncomp; k++) {
        HIP_TRY(hipMemcpy(tmp.data(), ptr + k * stride, nc * sizeof(float), hipMemcpyDeviceToHost));
        const double base = which == LFM_FIELD_Q ? qinf[k] : 0.0;
        double *o = out + k * nc;
        for (size_t ci = 0; ci < nc; ci++) o[L.orig[ci]] = base + (double)tmp[ci];
    }
    return LFM_OK;
}

void free_k32(KArgs32 *k) { delete k; }

int integrals(Solver *s, double *out_dev) {
    if (!s->fp32) return integrals64(s, out_dev);
    if (s->L.D == 2)
        launch_integrals32<2>(s, out_dev);
    else
        launch_integrals32<3>(s, out_dev);
    return 0;
}

int nonfinite32(Solver *s, unsigned long long *d) {
    if (s->L.D == 2) launch_nonfinite32<2>(s, d);
    else launch_nonfinite32<3>(s, d);
    return 0;
}

template void launch_bc32<2>(Solver *);
template void launch_bc32<3>(Solver *);
template void launch_gradvis32<2>(Solver *, bool, bool);
template void launch_gradvis32<3>(Solver *, bool, bool);
template void launch_flux32<2>(Solver *);
template void launch_flux32<3>(Solver *);
template void launch_update32<2>(Solver *, double, double, bool, bool, bool, const Solver::CflJob *);
template void launch_update32<3>(Solver *, double, double, bool, bool, bool, const Solver::CflJob *);
template void launch_stage_fused32<2>(Solver *, double, double, bool, const float *, float *,
                                     const Solver::CflJob *, bool);
template void launch_stage_fused32<3>(Solver *, double, double, bool, const float *, float *,
                                     const Solver::CflJob *, bool);
template void launch_build_slots32<2>(Solver *, int32_t *, float *, float *);
template void launch_build_slots32<3>(Solver *, int32_t *, float *, float *);
template void launch_cfl_dt32<2>(Solver *, int, double, double *, double *);
template void launch_cfl_dt32<3>(Solver *, int, double, double *, double *);

}  // namespace lfm
