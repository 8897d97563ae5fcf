// post.hip -- post-processing on the device: running pressure averages and
// wall forces, as Mesh::solve drives them between time steps.
//
// Reference behaviour restated:
//   postProcAverage   src/cfd_v0.cpp:3083-3103: for n > 0, p from the cell
//                     state, pAVG += p, pRMS += (p - pAVG/n)^2
//   updateAverageField src/cfd_v0.cpp:2966-2970 with polyMeshReaderOF.cpp:
//                     792-861: pAvg = pAVG * (1/n), pRMS = sqrt(pRMS * (1/n))
//   postProcForces    src/cfd_v0.cpp:3169-3233: per wall patch, pressure
//                     force sum S (p - p_inf) and viscous force -sum tau S
//                     with tau from the one-sided wall gradient
//   driver            src/mesh_solver.cpp:253-271, 423-429, 740-756: time +=
//                     dt (the new dt when adjusting), time_step++, average,
//                     forces every saveForcesStep steps
// The face each wall entry reads and the neighbour state it sees follow the
// reference: boundaries[patch][i] keeps the slot index the face had before
// reorder_faces while the cell's S, d and neighbour arrays are indexed after
// it, and a boundary ghost holds what the last set_boundary_conditions (the
// start of the step's last stage) put there.  The device keeps that ghost
// state in a snapshot taken before the last stage's update, because the
// update kernel refreshes the ghosts from the updated cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels_fp32p.hpp"
#include "kernels_fp64.hpp"
#include "lfm_hip.h"
#include "solver_state.hpp"

namespace lfm {

static constexpr int kPostBlock = 256;
static constexpr int kHistRows = 1024;  // force rows kept on the device before a drain

// State as stored (cell records): value = base[k] + (double) q(k, i) (base 0
// for the FP64 path, the free stream for the FP32 deviation path).
template <typename T>
struct QView {
    const T *q;
    int ntot;
    double base[5];
};

template <typename T>
__device__ __forceinline__ double qv(const QView<T> &v, int k, int i) {
    // cell records (components 0-3, then 4) in either precision
    return v.base[k] + (double)v.q[q64_index(k, (size_t)i, (size_t)v.ntot)];
}

template <int D, typename T>
__global__ void __launch_bounds__(kPostBlock) avg_accumulate(QView<T> v, int nc, double gm1, double n, double *pavg,
                                                             double *prms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double r = qv(v, 0, i);
    double velMag = 0.0;
#pragma unroll
    for (int k = 1; k <= D; k++) {
        const double vel = qv(v, k, i) / r;
        velMag += vel * vel;
    }
    const double E = qv(v, D + 1, i) / r;
    const double p = (r * gm1) * (E - 0.5 * velMag);
    const double a = pavg[i] + p;
    pavg[i] = a;
    const double dev = p - a / n;
    prms[i] += dev * dev;
}

template <typename T>
__global__ void __launch_bounds__(kPostBlock) ghost_snapshot(QView<T> v, const int32_t *idx, int n, int ncomp,
                                                             double *snap) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int g = idx[j];
    for (int k = 0; k < ncomp; k++) snap[(size_t)k * n + j] = qv(v, k, g);
}

__global__ void advance_time(double *time, const double *dt) {
    if (threadIdx.x == 0) time[0] += dt[0];
}

__global__ void write_time(double *row, const double *time_ptr, double time_val) {
    if (threadIdx.x == 0) row[0] = time_ptr ? time_ptr[0] : time_val;
}

// One block per wall patch; each thread sums a strided share of the patch's
// entries, then a fixed-shape tree adds the thread sums.
template <int D, typename T>
__global__ void __launch_bounds__(kPostBlock) wall_forces(QView<T> v, const int32_t *off, const int32_t *cell,
                                                          const int32_t *nb, const double *S, const double *d, int ne,
                                                          const double *snap, int nsnap, double gm1, double p_inf,
                                                          double mu, const double *time_ptr, double time_val,
                                                          double *row) {
    __shared__ double red[2 * D][kPostBlock];
    const int w = blockIdx.x, t = threadIdx.x;
    double acc[2 * D];
#pragma unroll
    for (int k = 0; k < 2 * D; k++) acc[k] = 0.0;
    for (int e = off[w] + t; e < off[w + 1]; e += kPostBlock) {
        const int c = cell[e];
        double q[D + 2], qn[D + 2], Sf[D], df[D];
#pragma unroll
        for (int k = 0; k < D + 2; k++) q[k] = qv(v, k, c);
        const int n = nb[e];
        if (n >= 0) {
#pragma unroll
            for (int k = 0; k < D + 2; k++) qn[k] = qv(v, k, n);
        } else {
#pragma unroll
            for (int k = 0; k < D + 2; k++) qn[k] = snap[(size_t)k * nsnap + ~n];
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
            Sf[k] = S[(size_t)k * ne + e];
            df[k] = d[(size_t)k * ne + e];
        }
        const double r = q[0];
        double Umag = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) {
            const double u = q[k + 1] / r;
            Umag += u * u;
        }
        Umag = sqrt(Umag);
        const double E = q[D + 1] / r;
        const double p = (r * gm1) * (E - 0.5 * Umag * Umag);
#pragma unroll
        for (int k = 0; k < D; k++) acc[k] += Sf[k] * (p - p_inf);
        double dm = df[0] * df[0];
#pragma unroll
        for (int k = 1; k < D; k++) dm += df[k] * df[k];
        const double dmag_inv = 1.0 / sqrt(dm);
        double dn[D], dudx[D][D], tau[D][D];
#pragma unroll
        for (int k = 0; k < D; k++) dn[k] = df[k] * dmag_inv;
#pragma unroll
        for (int i = 0; i < D; i++) {
            const double cu = q[i + 1] / q[0];
            const double au = qn[i + 1] / qn[0];
#pragma unroll
            for (int j = 0; j < D; j++) dudx[i][j] = (au - cu) * dn[j] * dmag_inv;
        }
#pragma unroll
        for (int a = 0; a < D; a++) {
            tau[a][a] = 2.0 * dudx[a][a];
#pragma unroll
            for (int a1 = a + 1; a1 < D + a; a1++) {
                const int a2 = a1 % D;
                tau[a][a2] = mu * (dudx[a][a2] + dudx[a2][a]);
                tau[a][a] -= dudx[a2][a2];
            }
            tau[a][a] *= 2.0 / 3.0 * mu;
        }
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++) acc[D + a] -= tau[a][b] * Sf[b];
    }
#pragma unroll
    for (int k = 0; k < 2 * D; k++) red[k][t] = acc[k];
    __syncthreads();
    for (int h = kPostBlock / 2; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int k = 0; k < 2 * D; k++) red[k][t] += red[k][t + h];
        __syncthreads();
    }
    if (t < 2 * D) row[1 + w * 2 * D + t] = red[t][0];
    if (w == 0 && t == 0) row[0] = time_ptr ? time_ptr[0] : time_val;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct PostState {
    lfm_post cfg{};
    int64_t step = 0;
    std::vector<void *> allocs;
    double *pavg = nullptr, *prms = nullptr;
    // wall forces
    std::vector<int32_t> walls;  // wall patch indices, patch order
    int ne = 0, nsnap = 0, row_w = 0, rows_dev = 0;
    int32_t *off = nullptr, *cell = nullptr, *nb = nullptr, *snap_idx = nullptr;
    double *S = nullptr, *d = nullptr, *snap = nullptr, *hist = nullptr;
    bool snap_valid = false;
    std::vector<double> rows_host;
};

void free_post(PostState *p) {
    if (!p) return;
    for (void *a : p->allocs) (void)hipFree(a);
    delete p;
}

template <typename T>
static int post_alloc(PostState *p, T **dst, size_t n, const T *src = nullptr) {
    *dst = nullptr;
    if (n == 0) return 0;
    HIP_TRY(hipMalloc((void **)dst, n * sizeof(T)));
    p->allocs.push_back((void *)*dst);
    if (src)
        HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    else
        HIP_TRY(hipMemset(*dst, 0, n * sizeof(T)));
    return 0;
}

template <typename T>
static QView<T> qview(Solver *s) {
    QView<T> v{};
    v.ntot = s->L.nc + s->L.ng;
    if constexpr (sizeof(T) == 4) {
        const KArgs32 &a = *s->k32;
        v.q = (const T *)a.q;
        v.base[0] = a.rho_inf_d;
        for (int k = 0; k < s->L.D; k++) v.base[k + 1] = a.m_inf_d[k];
        v.base[s->L.D + 1] = a.rhoE_inf_d;
    } else {
        v.q = (const T *)s->k64.q;
    }
    return v;
}

static const double *dt_ptr(Solver *s) { return s->fp32 ? s->k32->dt : s->k64.dt; }

void Solver::post_snapshot() {
    PostState *p = post;
    if (!p || !p->cfg.have_forces || p->nsnap == 0) {
        if (p) p->snap_valid = true;
        return;
    }
    const int nb = (p->nsnap + kPostBlock - 1) / kPostBlock;
    if (fp32)
        hipLaunchKernelGGL(ghost_snapshot<float>, dim3(nb), dim3(kPostBlock), 0, stream, qview<float>(this),
                           p->snap_idx, p->nsnap, L.D + 2, p->snap);
    else
        hipLaunchKernelGGL(ghost_snapshot<double>, dim3(nb), dim3(kPostBlock), 0, stream, qview<double>(this),
                           p->snap_idx, p->nsnap, L.D + 2, p->snap);
    p->snap_valid = true;
}

// records one force row (time from the device clock when time_ptr is set)
static int post_record(Solver *s, const double *time_ptr, double time_val) {
    PostState *p = s->post;
    if (s->ghosts_stale) {
        // the state was set from outside: the ghosts are filled as the
        // reference's start-up set_boundary_conditions does before its
        // first postProcForces (mesh_solver.cpp:416-429)
        s->bc();
        s->ghosts_stale = false;
        p->snap_valid = false;
    }
    if (!p->snap_valid) s->post_snapshot();
    if (p->rows_dev == kHistRows) {
        const size_t n0 = p->rows_host.size();
        p->rows_host.resize(n0 + (size_t)kHistRows * p->row_w);
        HIP_TRY(hipMemcpyAsync(p->rows_host.data() + n0, p->hist, (size_t)kHistRows * p->row_w * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        p->rows_dev = 0;
    }
    double *row = p->hist + (size_t)p->rows_dev * p->row_w;
    const int nw = (int)p->walls.size();
    const Gas &g = s->gas;
    if (nw == 0) {
        // no wall patch: the row is the time alone
        hipLaunchKernelGGL(write_time, dim3(1), dim3(64), 0, s->stream, row, time_ptr, time_val);
    } else if (s->L.D == 2) {
        if (s->fp32)
            hipLaunchKernelGGL((wall_forces<2, float>), dim3(nw), dim3(kPostBlock), 0, s->stream, qview<float>(s),
                               p->off, p->cell, p->nb, p->S, p->d, p->ne, p->snap, p->nsnap, g.gm1, g.p_inf, g.mu,
                               time_ptr, time_val, row);
        else
            hipLaunchKernelGGL((wall_forces<2, double>), dim3(nw), dim3(kPostBlock), 0, s->stream, qview<double>(s),
                               p->off, p->cell, p->nb, p->S, p->d, p->ne, p->snap, p->nsnap, g.gm1, g.p_inf, g.mu,
                               time_ptr, time_val, row);
    } else {
        if (s->fp32)
            hipLaunchKernelGGL((wall_forces<3, float>), dim3(nw), dim3(kPostBlock), 0, s->stream, qview<float>(s),
                               p->off, p->cell, p->nb, p->S, p->d, p->ne, p->snap, p->nsnap, g.gm1, g.p_inf, g.mu,
                               time_ptr, time_val, row);
        else
            hipLaunchKernelGGL((wall_forces<3, double>), dim3(nw), dim3(kPostBlock), 0, s->stream, qview<double>(s),
                               p->off, p->cell, p->nb, p->S, p->d, p->ne, p->snap, p->nsnap, g.gm1, g.p_inf, g.mu,
                               time_ptr, time_val, row);
    }
    p->rows_dev++;
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

static int post_average(Solver *s, int64_t n) {
    PostState *p = s->post;
    if (n <= 0) return LFM_OK;
    const int nc = s->L.nc;
    const int nb = (nc + kPostBlock - 1) / kPostBlock;
    const double gm1 = s->gas.gm1, dn = (double)n;
    if (s->L.D == 2) {
        if (s->fp32)
            hipLaunchKernelGGL((avg_accumulate<2, float>), dim3(nb), dim3(kPostBlock), 0, s->stream,
                               qview<float>(s), nc, gm1, dn, p->pavg, p->prms);
        else
            hipLaunchKernelGGL((avg_accumulate<2, double>), dim3(nb), dim3(kPostBlock), 0, s->stream,
                               qview<double>(s), nc, gm1, dn, p->pavg, p->prms);
    } else {
        if (s->fp32)
            hipLaunchKernelGGL((avg_accumulate<3, float>), dim3(nb), dim3(kPostBlock), 0, s->stream,
                               qview<float>(s), nc, gm1, dn, p->pavg, p->prms);
        else
            hipLaunchKernelGGL((avg_accumulate<3, double>), dim3(nb), dim3(kPostBlock), 0, s->stream,
                               qview<double>(s), nc, gm1, dn, p->pavg, p->prms);
    }
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int Solver::post_step() {
    PostState *p = post;
    if (!p) return LFM_OK;
    p->step++;
    if (p->cfg.have_forces)
        hipLaunchKernelGGL(advance_time, dim3(1), dim3(64), 0, stream, d_scalar + 15, dt_ptr(this));
    int rc = LFM_OK;
    if (p->cfg.have_average) rc = post_average(this, p->step - p->cfg.average_start);
    if (!rc && p->cfg.have_forces && p->step % p->cfg.forces_every == 0) rc = post_record(this, d_scalar + 15, 0.0);
    return rc;
}

static int post_configure(Solver *s, const lfm_post &cfg) {
    const SolverLayout &L = s->L;
    const int D = L.D;
    HIP_TRY(hipStreamSynchronize(s->stream));
    free_post(s->post);
    s->post = nullptr;
    s->post_forces = false;
    if (!cfg.have_average && !cfg.have_forces) return LFM_OK;
    auto *p = new PostState();
    s->post = p;
    p->cfg = cfg;
    int rc = 0;
    if (cfg.have_average) {
        rc |= post_alloc(p, &p->pavg, (size_t)L.nc);
        rc |= post_alloc(p, &p->prms, (size_t)L.nc);
    }
    if (cfg.have_forces) {
        for (size_t pch = 0; pch < L.patch_role.size(); pch++)
            if (L.patch_role[pch] == LFM_BC_WALL) p->walls.push_back((int32_t)pch);
        std::vector<int32_t> off(1, 0), cell, nb, snap_idx;
        std::vector<double> S, d;
        std::vector<int32_t> snap_of(L.ng_bnd, -1);
        std::vector<std::vector<double>> Sk(D), dk(D);
        for (int32_t pch : p->walls) {
            for (int g = 0; g < L.ng_bnd; g++) {
                if (L.ghost_patch[g] != pch) continue;
                cell.push_back(L.ghost_cell[g]);
                int n = cfg.forces_own_face ? L.nc + g : L.gf_ref_nb[g];
                const std::vector<double> &Sv = cfg.forces_own_face ? L.gf_own_S : L.gf_ref_S;
                const std::vector<double> &dv = cfg.forces_own_face ? L.gf_own_d : L.gf_ref_d;
                for (int k = 0; k < D; k++) {
                    Sk[k].push_back(Sv[(size_t)k * L.ng_bnd + g]);
                    dk[k].push_back(dv[(size_t)k * L.ng_bnd + g]);
                }
                if (n >= L.nc && n < L.nc + L.ng_bnd) {
                    // a boundary ghost: read from the snapshot
                    int &sl = snap_of[n - L.nc];
                    if (sl < 0) {
                        sl = (int)snap_idx.size();
                        snap_idx.push_back(n);
                    }
                    n = ~sl;
                }
                nb.push_back(n);
            }
            off.push_back((int32_t)cell.size());
        }
        p->ne = (int)cell.size();
        p->nsnap = (int)snap_idx.size();
        for (int k = 0; k < D; k++) {
            S.insert(S.end(), Sk[k].begin(), Sk[k].end());
            d.insert(d.end(), dk[k].begin(), dk[k].end());
        }
        p->row_w = 1 + 2 * D * (int)p->walls.size();
        rc |= post_alloc(p, &p->off, off.size(), off.data());
        rc |= post_alloc(p, &p->cell, cell.size(), cell.data());
        rc |= post_alloc(p, &p->nb, nb.size(), nb.data());
        rc |= post_alloc(p, &p->S, S.size(), S.data());
        rc |= post_alloc(p, &p->d, d.size(), d.data());
        rc |= post_alloc(p, &p->snap_idx, snap_idx.size(), snap_idx.data());
        rc |= post_alloc(p, &p->snap, (size_t)(D + 2) * p->nsnap);
        rc |= post_alloc(p, &p->hist, (size_t)kHistRows * p->row_w);
        if (!rc) HIP_TRY(hipMemcpy(s->d_scalar + 15, &cfg.time, sizeof(double), hipMemcpyHostToDevice));
        s->post_forces = true;
    }
    if (rc) {
        free_post(s->post);
        s->post = nullptr;
        s->post_forces = false;
        return rc < 0 ? rc : LFM_EDEVICE;
    }
    return LFM_OK;
}

}  // namespace lfm

using lfm::fail;
using lfm::Solver;

extern "C" {

int lfm_post_configure(lfm_solver *h, const lfm_post *cfg) {
    if (!h || !cfg) return fail(LFM_EINVAL, "null argument");
    if (cfg->have_forces && cfg->forces_every < 1) return fail(LFM_EINVAL, "forces_every must be >= 1");
    return lfm::post_configure(h->impl.get(), *cfg);
}

int lfm_post_average(lfm_solver *h, int32_t n) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    Solver *s = h->impl.get();
    if (!s->post || !s->post->cfg.have_average) return fail(LFM_ESTATE, "averages not configured");
    return lfm::post_average(s, n);
}

int lfm_post_average_fields(lfm_solver *h, int32_t n, double *pavg, double *prms) {
    if (!h || !pavg || !prms) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    if (!s->post || !s->post->cfg.have_average) return fail(LFM_ESTATE, "averages not configured");
    if (n <= 0) return fail(LFM_EINVAL, "n must be positive");
    const size_t nc = s->L.nc;
    std::vector<double> a(nc), r(nc);
    HIP_TRY(hipMemcpyAsync(a.data(), s->post->pavg, nc * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(r.data(), s->post->prms, nc * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const double f = 1.0 / n;
    for (size_t ci = 0; ci < nc; ci++) {
        pavg[s->L.orig[ci]] = a[ci] * f;
        prms[s->L.orig[ci]] = std::sqrt(r[ci] * f);
    }
    return LFM_OK;
}

int lfm_post_walls(lfm_solver *h, int32_t *n, int32_t *patches) {
    if (!h || !n) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    if (!s->post || !s->post->cfg.have_forces) return fail(LFM_ESTATE, "forces not configured");
    *n = (int32_t)s->post->walls.size();
    if (patches) std::copy(s->post->walls.begin(), s->post->walls.end(), patches);
    return LFM_OK;
}

int lfm_post_forces(lfm_solver *h, double time) {
    if (!h) return fail(LFM_EINVAL, "null handle");
    Solver *s = h->impl.get();
    if (!s->post || !s->post->cfg.have_forces) return fail(LFM_ESTATE, "forces not configured");
    return lfm::post_record(s, nullptr, time);
}

int lfm_post_forces_read(lfm_solver *h, int32_t max_rows, double *rows, int32_t *n_rows) {
    if (!h || !n_rows || max_rows < 0) return fail(LFM_EINVAL, "bad argument");
    Solver *s = h->impl.get();
    lfm::PostState *p = s->post;
    if (!p || !p->cfg.have_forces) return fail(LFM_ESTATE, "forces not configured");
    if (p->rows_dev) {
        const size_t n0 = p->rows_host.size();
        p->rows_host.resize(n0 + (size_t)p->rows_dev * p->row_w);
        HIP_TRY(hipMemcpyAsync(p->rows_host.data() + n0, p->hist, (size_t)p->rows_dev * p->row_w * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        p->rows_dev = 0;
    }
    const int avail = (int)(p->rows_host.size() / p->row_w);
    if (!rows) {
        *n_rows = avail;
        return LFM_OK;
    }
    const int n = std::min(avail, (int)max_rows);
    std::copy(p->rows_host.begin(), p->rows_host.begin() + (size_t)n * p->row_w, rows);
    p->rows_host.erase(p->rows_host.begin(), p->rows_host.begin() + (size_t)n * p->row_w);
    *n_rows = n;
    return LFM_OK;
}

}  // extern "C"
