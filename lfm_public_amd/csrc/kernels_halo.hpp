// kernels_halo.hpp -- halo pack / unpack between partitions.
//
// A payload is a sequence of per-cell records in the reference's send-buffer
// order (cfd_v0.cpp:3303-3331 packed, 3398-3408 solvars, 3475-3499 viscous):
//   SOLVARS: q[0..D+1]
//   VISCOUS: dudx[i][j] (row-major), dTdx[i], tauMC[i][j], sigmaU[i]
//   PACKED : SOLVARS then VISCOUS
// The state and the gradients are stored as per-cell records, tauMC and
// sigmaU as [component][ntot]: one thread per (record, component) keeps the
// buffer writes (pack) and reads (unpack) contiguous.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lfm {

template <typename T>
struct HaloFields {
    T *q, *dudx, *dTdx, *tauMC, *sigmaU;
    int ntot, D, kind;  // kind: bit 0 SOLVARS, bit 1 VISCOUS
    int q_records;      // q stored as cell records (components 0-3, then 4)
    int grad_stride;    // dudx / dTdx in per-cell records of this stride, 0: SoA
};

// index of component m of cell i in field f
template <typename T>
__device__ __forceinline__ size_t halo_index(const HaloFields<T> &h, const T *f, int m, int i) {
    if (f == h.q && h.q_records) return m < 4 ? 4 * (size_t)i + m : 4 * (size_t)h.ntot + i;
    if ((f == h.dudx || f == h.dTdx) && h.grad_stride) return (size_t)i * h.grad_stride + m;
    return (size_t)m * h.ntot + i;
}

template <typename T>
__device__ __forceinline__ T *halo_component(const HaloFields<T> &h, int k, int &m) {
    const int D = h.D;
    if (h.kind & 1) {
        if (k < D + 2) {
            m = k;
            return h.q;
        }
        k -= D + 2;
    }
    if (k < D * D) {
        m = k;
        return h.dudx;
    }
    k -= D * D;
    if (k < D) {
        m = k;
        return h.dTdx;
    }
    k -= D;
    if (k < D * D) {
        m = k;
        return h.tauMC;
    }
    m = k - D * D;
    return h.sigmaU;
}

// buf[i * nrec + k] = component k of cell cells[i]
template <typename T>
__global__ void __launch_bounds__(256) halo_pack_kernel(HaloFields<T> h, const int32_t *cells, int n, int nrec, T *buf) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * nrec) return;
    const int i = (int)(t / nrec), k = (int)(t % nrec);
    int m;
    const T *src = halo_component(h, k, m);
    buf[t] = src[halo_index(h, src, m, cells[i])];
}

// component k of ghost slot ghost[j] = buf[src[j] * nrec + k]
template <typename T>
__global__ void __launch_bounds__(256) halo_unpack_kernel(HaloFields<T> h, const int32_t *ghost, const int32_t *src, int n,
                                                   int nrec, const T *buf) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * nrec) return;
    const int j = (int)(t / nrec), k = (int)(t % nrec);
    int m;
    T *dst = halo_component(h, k, m);
    dst[halo_index(h, dst, m, ghost[j])] = buf[(size_t)src[j] * nrec + k];
}

// dt = min over partitions of their dt minima (MPI_Allreduce MIN,
// mesh_solver.cpp:705-717), written to every partition's deltaT
static __global__ void group_dt_min(double *const *dtmin, double *const *dt, int n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v = *dtmin[0];
    for (int i = 1; i < n; i++) {
        const double o = *dtmin[i];
        if (v > o) v = o;
    }
    for (int i = 0; i < n; i++) *dt[i] = v;
}

}  // namespace lfm
