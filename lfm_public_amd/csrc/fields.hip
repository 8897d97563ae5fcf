// fields.hip -- host <-> device field transfer in original cell order, state
// reset for a new time step, and the finiteness check.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
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

// ---------------------------------------------------------------------------
// The finiteness count without allocating (lfm_check_finite_noalloc): per-
// block counts in a file-scope device array, one slot per block of 256
// cells, the total in the slot after them.  Every block writes its own slot
// and the final kernel writes the total, so nothing needs clearing between
// calls.  One array per device, shared by every call on it: calls that
// overlap on one device race.
// ---------------------------------------------------------------------------
constexpr int kNoAllocBlocks = 65536;  // cells <= 65536 * 256 = 16,777,216
static __device__ __align__(256) unsigned long long scratch[kNoAllocBlocks + 1];

// the array's device address, looked up once per device with that device
// current (the host-side name is not a device address)
static std::mutex g_scratch_mutex;
static std::vector<unsigned long long *> g_scratch_ptr;

static int scratch_of(int device, unsigned long long **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = (unsigned long long *)p;
    }
    *out = g_scratch_ptr[device];
    return LFM_OK;
}

// per-block count of non-finite conservatives (cells c < nc of a state in
// the record layout: components 0-3 at q[4 i + k], the 3D energy at
// q[4 ntot + i]); every block writes partial[blockIdx.x]
template <typename T, int D>
static __global__ void __launch_bounds__(256) nonfinite_blocks(const T *q, int nc, size_t ntot,
                                                               unsigned long long *partial) {
    __shared__ unsigned long long red[256 / 64];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long bad = 0;
    if (c < nc) {
#pragma unroll
        for (int k = 0; k < D + 2; k++) bad += !isfinite(q[k < 4 ? 4 * (size_t)c + k : 4 * ntot + c]);
    }
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bad;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// the sum of nb block counts into partial[total_slot] (one block of 1024)
static __global__ void __launch_bounds__(1024) nonfinite_total(unsigned long long *partial, int nb, int total_slot) {
    __shared__ unsigned long long red[1024 / 64];
    unsigned long long v = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v += partial[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < 1024 / 64; i++) t += red[i];
        partial[total_slot] = t;
    }
}

int Solver::check_finite_noalloc(int64_t *n_bad) {
    const int nb = (L.nc + 255) / 256;
    if (nb > kNoAllocBlocks)
        return fail(LFM_EINVAL, "lfm_check_finite_noalloc: " + std::to_string(L.nc) + " cells exceed its maximum of " +
                                    std::to_string((long)kNoAllocBlocks * 256) + " (use lfm_check_finite)");
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    unsigned long long *part = nullptr;
    int rc = scratch_of(device, &part);
    if (rc == LFM_OK && nb > 0) {
        const size_t ntot = (size_t)L.nc + L.ng;
        if (fp32) {
            if (L.D == 2)
                hipLaunchKernelGGL((nonfinite_blocks<float, 2>), dim3(nb), dim3(256), 0, stream, k32->q, L.nc, ntot, part);
            else
                hipLaunchKernelGGL((nonfinite_blocks<float, 3>), dim3(nb), dim3(256), 0, stream, k32->q, L.nc, ntot, part);
        } else {
            if (L.D == 2)
                hipLaunchKernelGGL((nonfinite_blocks<double, 2>), dim3(nb), dim3(256), 0, stream, k64.q, L.nc, ntot, part);
            else
                hipLaunchKernelGGL((nonfinite_blocks<double, 3>), dim3(nb), dim3(256), 0, stream, k64.q, L.nc, ntot, part);
        }
        hipLaunchKernelGGL(nonfinite_total, dim3(1), dim3(1024), 0, stream, part, nb, kNoAllocBlocks);
        hipError_t e = hipGetLastError();
        // the total into the handle's pinned host scratch
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_scalar + 40, part + kNoAllocBlocks, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) rc = fail(LFM_EDEVICE, std::string("lfm_check_finite_noalloc: ") + hipGetErrorString(e));
    }
    if (rc == LFM_OK) {
        unsigned long long t = 0;
        if (nb > 0) std::memcpy(&t, h_scalar + 40, sizeof t);
        *n_bad = (int64_t)t;
    }
    if (cur != device) (void)hipSetDevice(cur);
    return rc;
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
