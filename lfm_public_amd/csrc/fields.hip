// fields.hip -- host <-> device field transfer in original cell order, state
// reset for a new time step, and the finiteness check.
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels_fp32p.hpp"
#include "solver_state.hpp"

namespace lfm {

// ISolver::prepare_for_timestep: dq = 0, RES = 0
static __global__ void __launch_bounds__(256) fill_zero32(float *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

void Solver::zero_state_for_step() {
    const size_t n = (size_t)(L.D + 2) * L.nc;
    const int nb = (int)((n + 255) / 256);
    if (fp32) {
        hipLaunchKernelGGL(fill_zero32, dim3(nb), dim3(256), 0, stream, k32->dq, n);
        hipLaunchKernelGGL(fill_zero32, dim3(nb), dim3(256), 0, stream, k32->res_cell, n);
    } else {
        hipLaunchKernelGGL(fill_zero, dim3(nb), dim3(256), 0, stream, k64.dq, n);
        hipLaunchKernelGGL(fill_zero, dim3(nb), dim3(256), 0, stream, k64.res_cell, n);
    }
    pending_A = 1.0;
}

namespace {

// A field in the FP64 layout: component k of cell ci at ptr[k * cs + ci * es],
// or, for the state (q), at ptr[q64_index(k, ci, ntot)]
struct FieldRef {
    double *ptr;
    int ncomp;
    size_t cs, es;
    bool state;
};

FieldRef field_ref64(Solver *s, int which) {
    const int D = s->L.D;
    const size_t nt = (size_t)s->L.nc + s->L.ng, nc = s->L.nc, gs = (size_t)s->k64.gstride;
    KArgs &a = s->k64;
    switch (which) {
    case LFM_FIELD_Q: return {a.q, D + 2, 0, 0, true};
    case LFM_FIELD_DQ: return {a.dq, D + 2, nc, 1, false};
    case LFM_FIELD_DUDX: return {a.dudx, D * D, 1, gs, false};
    case LFM_FIELD_DTDX: return {a.dTdx, D, 1, gs, false};
    case LFM_FIELD_TAUMC: return {a.tauMC, D * D, nt, 1, false};
    case LFM_FIELD_SIGMAU: return {a.sigmaU, D, nt, 1, false};
    case LFM_FIELD_RHOGRAD: return {a.rho_grad, D, 1, gs, false};
    case LFM_FIELD_PGRAD: return {a.p_grad, D, 1, gs, false};
    case LFM_FIELD_UGRAD: return {a.U_grad, D * D, 1, gs, false};
    case LFM_FIELD_RES: return {a.res_cell, D + 2, nc, 1, false};
    default: return {nullptr, 0, 0, 0, false};
    }
}

}  // namespace

int Solver::get_field(int which, double *out) {
    const int D = L.D;
    const size_t nc = L.nc;
    HIP_TRY(hipStreamSynchronize(stream));
    if (which == LFM_FIELD_PRIM) {
        std::vector<double> q((size_t)(D + 2) * nc);
        int rc = get_field(LFM_FIELD_Q, q.data());
        if (rc) return rc;
        for (size_t c = 0; c < nc; c++) {
            const double rho = q[c];
            double u2 = 0.0;
            for (int k = 0; k < D; k++) {
                const double u = q[(k + 1) * nc + c] / rho;
                out[(k + 1) * nc + c] = u;
                u2 += u * u;
            }
            const double p = (q[(D + 1) * nc + c] - 0.5 * rho * u2) * gas.gm1;
            out[c] = p;
            out[(D + 1) * nc + c] = p * gas.R_inv / rho;
        }
        return LFM_OK;
    }
    if (fp32) return get_field32(this, which, out);
    FieldRef r = field_ref64(this, which);
    if (!r.ptr && (which == LFM_FIELD_TAUMC || which == LFM_FIELD_SIGMAU)) {
        // not stored: formed from dudx and the state, as calc_VIS leaves them
        double *dev = nullptr;
        const size_t n = (size_t)(D * D + D) * nc;
        HIP_TRY(hipMalloc((void **)&dev, n * sizeof(double)));
        const int nb = (int)((nc + 255) / 256);
        if (D == 2)
            hipLaunchKernelGGL(vis_tensor_field<2>, dim3(nb), dim3(256), 0, stream, k64, dev, dev + (size_t)D * D * nc);
        else
            hipLaunchKernelGGL(vis_tensor_field<3>, dim3(nb), dim3(256), 0, stream, k64, dev, dev + (size_t)D * D * nc);
        std::vector<double> tmp(n);
        hipError_t e = hipMemcpyAsync(tmp.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        (void)hipFree(dev);
        HIP_TRY(e);
        const int ncomp = which == LFM_FIELD_TAUMC ? D * D : D;
        const double *src = tmp.data() + (which == LFM_FIELD_TAUMC ? 0 : (size_t)D * D * nc);
        for (int k = 0; k < ncomp; k++)
            for (size_t ci = 0; ci < nc; ci++) out[k * nc + L.orig[ci]] = src[k * nc + ci];
        return LFM_OK;
    }
    if (!r.ptr) return fail(LFM_EINVAL, "unknown or unallocated field");
    const size_t nt = (size_t)L.nc + L.ng;
    if (r.state) {
        std::vector<double> tmp((size_t)(D + 2) * nt);
        HIP_TRY(hipMemcpy(tmp.data(), r.ptr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < r.ncomp; k++)
            for (size_t ci = 0; ci < nc; ci++) out[k * nc + L.orig[ci]] = tmp[q64_index(k, ci, nt)];
    } else if (r.es == 1) {
        std::vector<double> tmp(nc);
        for (int k = 0; k < r.ncomp; k++) {
            HIP_TRY(hipMemcpy(tmp.data(), r.ptr + k * r.cs, nc * sizeof(double), hipMemcpyDeviceToHost));
            double *o = out + k * nc;
            for (size_t ci = 0; ci < nc; ci++) o[L.orig[ci]] = tmp[ci];
        }
    } else {
        // per-cell records: one copy of the cells' records from the field's slot
        std::vector<double> tmp((nc - 1) * r.es + r.ncomp);
        HIP_TRY(hipMemcpy(tmp.data(), r.ptr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < r.ncomp; k++)
            for (size_t ci = 0; ci < nc; ci++) out[k * nc + L.orig[ci]] = tmp[ci * r.es + k * r.cs];
    }
    return LFM_OK;
}

int Solver::set_field(int which, const double *in) {
    const size_t nc = L.nc;
    HIP_TRY(hipStreamSynchronize(stream));
    ghosts_stale = true;
    if (fp32) {
        if (which != LFM_FIELD_Q) return fail(LFM_EINVAL, "only q is settable on the FP32 path");
        KArgs32 &a = *k32;
        double qinf[5] = {a.rho_inf_d, a.m_inf_d[0], a.m_inf_d[1], a.m_inf_d[2], 0.0};
        qinf[L.D + 1] = a.rhoE_inf_d;
        // cell records: components 0-3 interleaved, component 4 after them
        const size_t nt = (size_t)L.nc + L.ng;
        std::vector<float> rec((size_t)4 * nc), e4(L.D == 3 ? nc : 0);
        for (int k = 0; k < L.D + 2; k++) {
            const double *src = in + k * nc;
            for (size_t ci = 0; ci < nc; ci++) {
                const float v = (float)(src[L.orig[ci]] - qinf[k]);
                if (k < 4)
                    rec[4 * ci + k] = v;
                else
                    e4[ci] = v;
            }
        }
        HIP_TRY(hipMemcpy(a.q, rec.data(), rec.size() * sizeof(float), hipMemcpyHostToDevice));
        if (L.D == 3) HIP_TRY(hipMemcpy(a.q + 4 * nt, e4.data(), nc * sizeof(float), hipMemcpyHostToDevice));
        return LFM_OK;
    }
    FieldRef r = field_ref64(this, which);
    if (!r.ptr || (which != LFM_FIELD_Q && which != LFM_FIELD_DQ)) return fail(LFM_EINVAL, "field not settable");
    if (r.state) {
        // the cells' records (ghost slots kept)
        const size_t nt = (size_t)L.nc + L.ng;
        std::vector<double> tmp((size_t)(L.D + 2) * nt);
        HIP_TRY(hipMemcpy(tmp.data(), r.ptr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < r.ncomp; k++)
            for (size_t ci = 0; ci < nc; ci++) tmp[q64_index(k, ci, nt)] = in[k * nc + L.orig[ci]];
        HIP_TRY(hipMemcpy(r.ptr, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
        return LFM_OK;
    }
    std::vector<double> tmp(nc);
    for (int k = 0; k < r.ncomp; k++) {
        const double *src = in + k * nc;
        for (size_t ci = 0; ci < nc; ci++) tmp[ci] = src[L.orig[ci]];
        HIP_TRY(hipMemcpy(r.ptr + k * r.cs, tmp.data(), nc * sizeof(double), hipMemcpyHostToDevice));
    }
    return LFM_OK;
}

int Solver::check_finite(int64_t *n_bad) {
    unsigned long long *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(d, 0, sizeof(unsigned long long), stream));
    const int nb = (L.nc + 255) / 256;
    if (fp32) {
        nonfinite32(this, d);
    } else {
        if (L.D == 2)
            hipLaunchKernelGGL(count_nonfinite<2>, dim3(nb), dim3(256), 0, stream, k64.q, L.nc, L.nc + L.ng, d);
        else
            hipLaunchKernelGGL(count_nonfinite<3>, dim3(nb), dim3(256), 0, stream, k64.q, L.nc, L.nc + L.ng, d);
    }
    unsigned long long h = 0;
    HIP_TRY(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipFree(d));
    *n_bad = (int64_t)h;
    return LFM_OK;
}

}  // namespace lfm
