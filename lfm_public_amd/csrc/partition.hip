// partition.hip -- halo exchange between partitions and the one-device
// partition group (ISolver::mpi_communication / mpi_wait and the MPI set-up
// of the reference, run between resident partitions on one GPU).
//
// Reference behaviour restated:
//   set-up messages   src/mesh_reader.cpp:512-612 (face id, tag, owner cell,
//                     cell-centre displacement), src/cfd_v0.cpp:616-700
//                     (send lists: boundary-submesh cells adjacent to the
//                     neighbour, in cell order)
//   payloads          src/cfd_v0.cpp:3303-3331, 3398-3408, 3475-3499 (pack),
//                     3608-3692 (unpack into ghost_mpi[neighbour][remote id])
//   stage sequencing  src/mesh_solver.cpp:409-419, 496-700: solvars exchanged
//                     after the RK update, viscous tensors after calc_VIS,
//                     dt reduced with MPI_Allreduce(MIN) when adjusting.
// Halo ghosts are one per processor face, holding the record of the
// neighbour cell that owns the face; the reference keeps one ghost per
// neighbour cell and points the faces at it.  Both read the same values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "kernels_fp32p.hpp"
#include "kernels_fp64.hpp"
#include "kernels_halo.hpp"
#include "lfm_hip.h"
#include "solver_state.hpp"

namespace lfm {

static constexpr int kBlock = 256;
static inline int nblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
static HaloFields<T> fields_of(T *q, T *dudx, T *dTdx, T *tau, T *sU, int ntot, int D, int kind, int grad_stride) {
    HaloFields<T> h;
    h.q = q;
    h.dudx = dudx;
    h.dTdx = dTdx;
    h.tauMC = tau;
    h.sigmaU = sU;
    h.ntot = ntot;
    h.D = D;
    h.kind = kind;
    h.q_records = 1;  // the state is stored as cell records
    h.grad_stride = grad_stride;
    return h;
}

int Solver::halo_pack(int ni, int kind, void *buf) {
    const HaloDev &hd = halo_dev[ni];
    const int rec = halo_record(kind);
    const long n = (long)hd.n_send * rec;
    if (n == 0) return LFM_OK;
    const int ntot = L.nc + L.ng;
    if (fp32) {
        KArgs32 &a = *k32;
        hipLaunchKernelGGL(halo_pack_kernel<float>, dim3(nblocks(n)), dim3(kBlock), 0, stream,
                           fields_of<float>(a.q, a.dudx, a.dTdx, a.tauMC, a.sigmaU, ntot, L.D, kind, a.gstride), hd.send_cells,
                           hd.n_send, rec, (float *)buf);
    } else {
        KArgs &a = k64;
        hipLaunchKernelGGL(halo_pack_kernel<double>, dim3(nblocks(n)), dim3(kBlock), 0, stream,
                           fields_of<double>(a.q, a.dudx, a.dTdx, a.tauMC, a.sigmaU, ntot, L.D, kind, a.gstride), hd.send_cells,
                           hd.n_send, rec, (double *)buf);
    }
    return LFM_OK;
}

int Solver::halo_unpack(int ni, int kind, const void *buf) {
    const HaloDev &hd = halo_dev[ni];
    const int rec = halo_record(kind);
    const long n = (long)hd.n_recv * rec;
    if (n == 0) return LFM_OK;
    const int ntot = L.nc + L.ng;
    if (fp32) {
        KArgs32 &a = *k32;
        hipLaunchKernelGGL(halo_unpack_kernel<float>, dim3(nblocks(n)), dim3(kBlock), 0, stream,
                           fields_of<float>(a.q, a.dudx, a.dTdx, a.tauMC, a.sigmaU, ntot, L.D, kind, a.gstride), hd.recv_ghost,
                           hd.recv_src, hd.n_recv, rec, (const float *)buf);
    } else {
        KArgs &a = k64;
        hipLaunchKernelGGL(halo_unpack_kernel<double>, dim3(nblocks(n)), dim3(kBlock), 0, stream,
                           fields_of<double>(a.q, a.dudx, a.dTdx, a.tauMC, a.sigmaU, ntot, L.D, kind, a.gstride), hd.recv_ghost,
                           hd.recv_src, hd.n_recv, rec, (const double *)buf);
    }
    return LFM_OK;
}

int Solver::halo_upload_geometry() {
    const size_t nf = L.nf, D = L.D;
    HIP_TRY(hipStreamSynchronize(stream));
    if (fp32) {
        std::vector<float> d(L.d.size()), w(L.w.size());
        for (size_t i = 0; i < d.size(); i++) d[i] = (float)L.d[i];
        for (size_t i = 0; i < w.size(); i++) w[i] = (float)L.w[i];
        HIP_TRY(hipMemcpy((void *)k32->d, d.data(), D * nf * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy((void *)k32->w, w.data(), nf * sizeof(float), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy((void *)k64.d, L.d.data(), D * nf * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy((void *)k64.w, L.w.data(), nf * sizeof(double), hipMemcpyHostToDevice));
    }
    return LFM_OK;
}

// Largest record index a neighbour's buffer must hold for this partition.
static int max_recv_src(const HaloNbr &h) {
    int m = -1;
    for (int v : h.recv_src) m = std::max(m, v);
    return m;
}

}  // namespace lfm

using lfm::fail;
using lfm::Solver;

struct lfm_part_group {
    std::vector<Solver *> parts;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // per part, per neighbour: send buffers of the two payloads
    std::vector<std::vector<void *>> buf_s, buf_v;
    // per part q, per neighbour i: index of q among that neighbour's neighbours
    std::vector<std::vector<int>> peer;
    double **d_dtmin = nullptr, **d_dt = nullptr;
    bool adjust = false;
    std::vector<void *> allocs;
    ~lfm_part_group() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int lfm_halo_neighbours(lfm_solver *h, int32_t *n, int32_t *ranks) {
    if (!h || !n) return fail(LFM_EINVAL, "null argument");
    const Solver *s = h->impl.get();
    *n = (int32_t)s->L.halo.size();
    if (ranks)
        for (size_t i = 0; i < s->L.halo.size(); i++) ranks[i] = s->L.halo[i].rank;
    return LFM_OK;
}

int lfm_halo_export(lfm_solver *h, int32_t rank, int32_t *n_faces, int32_t *keys, int32_t *cells, double *disp) {
    if (!h || !n_faces) return fail(LFM_EINVAL, "null argument");
    const Solver *s = h->impl.get();
    const int ni = s->halo_index(rank);
    if (ni < 0) return fail(LFM_EINVAL, "rank is not a neighbour of this partition");
    const lfm::HaloNbr &hn = s->L.halo[ni];
    const int n = (int)hn.key.size();
    *n_faces = n;
    for (int j = 0; j < n; j++) {
        if (keys) keys[j] = hn.key[j];
        if (cells) cells[j] = hn.cell[j];
        if (disp)
            for (int k = 0; k < 3; k++) disp[3 * j + k] = hn.xc[3 * j + k] - hn.xf[3 * j + k];
    }
    return LFM_OK;
}

int lfm_halo_import(lfm_solver *h, int32_t rank, int32_t n_faces, const int32_t *keys, const int32_t *cells,
                    const double *disp) {
    if (!h || (n_faces > 0 && (!keys || !cells || !disp))) return fail(LFM_EINVAL, "null argument");
    Solver *s = h->impl.get();
    const int ni = s->halo_index(rank);
    if (ni < 0) return fail(LFM_EINVAL, "rank is not a neighbour of this partition");
    for (int j = 0; j < n_faces; j++)
        if (cells[j] < 0) return fail(LFM_EINVAL, "negative cell index in the set-up message");
    std::string err;
    try {
        err = lfm::halo_import(s->L, ni, n_faces, keys, cells, disp);
    } catch (const std::exception &e) {
        return fail(LFM_ENOMEM, std::string("applying the set-up message: ") + e.what());
    }
    if (!err.empty()) return fail(LFM_EINVAL, err);
    std::lock_guard<std::mutex> lock(lfm::device_setup_mutex());
    int rc = s->halo_upload_geometry();
    if (rc) return rc;
    const lfm::HaloNbr &hn = s->L.halo[ni];
    HIP_TRY(hipMemcpy(s->halo_dev[ni].recv_src, hn.recv_src.data(), hn.recv_src.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
    return LFM_OK;
}

int lfm_halo_sizes(lfm_solver *h, int32_t rank, int kind, int64_t *send_bytes, int64_t *recv_bytes) {
    if (!h || !send_bytes || !recv_bytes) return fail(LFM_EINVAL, "null argument");
    if (kind < 1 || kind > 3) return fail(LFM_EINVAL, "payload must be SOLVARS, VISCOUS or PACKED");
    const Solver *s = h->impl.get();
    const int ni = s->halo_index(rank);
    if (ni < 0) return fail(LFM_EINVAL, "rank is not a neighbour of this partition");
    const int64_t el = s->fp32 ? 4 : 8, rec = s->halo_record(kind);
    *send_bytes = (int64_t)s->halo_dev[ni].n_send * rec * el;
    *recv_bytes = s->L.halo[ni].imported ? (int64_t)(lfm::max_recv_src(s->L.halo[ni]) + 1) * rec * el : 0;
    return LFM_OK;
}

int lfm_halo_pack(lfm_solver *h, int32_t rank, int kind, void *dev_buf) {
    if (!h || !dev_buf) return fail(LFM_EINVAL, "null argument");
    if (kind < 1 || kind > 3) return fail(LFM_EINVAL, "payload must be SOLVARS, VISCOUS or PACKED");
    Solver *s = h->impl.get();
    const int ni = s->halo_index(rank);
    if (ni < 0) return fail(LFM_EINVAL, "rank is not a neighbour of this partition");
    int rc = s->halo_pack(ni, kind, dev_buf);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_halo_unpack(lfm_solver *h, int32_t rank, int kind, const void *dev_buf) {
    if (!h || !dev_buf) return fail(LFM_EINVAL, "null argument");
    if (kind < 1 || kind > 3) return fail(LFM_EINVAL, "payload must be SOLVARS, VISCOUS or PACKED");
    Solver *s = h->impl.get();
    const int ni = s->halo_index(rank);
    if (ni < 0) return fail(LFM_EINVAL, "rank is not a neighbour of this partition");
    if (!s->L.halo[ni].imported) return fail(LFM_ESTATE, "halo set-up of this neighbour not imported");
    int rc = s->halo_unpack(ni, kind, dev_buf);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_part_group_create(lfm_solver *const *hs, int n, lfm_part_group **out) {
    if (!hs || !out || n < 1) return fail(LFM_EINVAL, "bad argument");
    *out = nullptr;
    std::vector<Solver *> parts(n);
    for (int i = 0; i < n; i++) {
        if (!hs[i] || !hs[i]->impl) return fail(LFM_EINVAL, "null partition");
        parts[i] = hs[i]->impl.get();
        if (parts[i]->L.rank != i && !(n == 1 && parts[i]->L.rank < 0))
            return fail(LFM_EINVAL, "parts[i] must be the partition of rank i");
        if (!parts[i]->halo_ready()) return fail(LFM_ESTATE, "halo set-up not imported for every neighbour");
    }
    const Solver *p0 = parts[0];
    for (Solver *p : parts)
        if (p->fp32 != p0->fp32 || p->L.D != p0->L.D || p->prm.solver != p0->prm.solver ||
            p->prm.minmod != p0->prm.minmod || p->nstages != p0->nstages || p->prm.adjust_dt != p0->prm.adjust_dt)
            return fail(LFM_EINVAL, "partitions must share precision, dimension, scheme, rkOrder and adjustTimeStep");
    auto g = std::make_unique<lfm_part_group>();
    g->parts = parts;
    g->adjust = p0->prm.adjust_dt != 0;
    g->peer.resize(n);
    for (int q = 0; q < n; q++) {
        const auto &hq = parts[q]->L.halo;
        for (size_t i = 0; i < hq.size(); i++) {
            const int r = hq[i].rank;
            if (r < 0 || r >= n || r == q) return fail(LFM_EINVAL, "neighbour rank outside the group");
            const int j = parts[r]->halo_index(q);
            if (j < 0) return fail(LFM_EINVAL, "processor patches are not mutual");
            // every record this side reads must exist in the sender's buffer
            if (lfm::max_recv_src(hq[i]) >= parts[r]->halo_dev[j].n_send)
                return fail(LFM_EINVAL, "receive list exceeds the neighbour's send list");
            g->peer[q].push_back(j);
        }
    }
    std::lock_guard<std::mutex> lock(lfm::device_setup_mutex());
    HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
    const size_t el = p0->fp32 ? 4 : 8;
    g->buf_s.resize(n);
    g->buf_v.resize(n);
    for (int p = 0; p < n; p++) {
        for (size_t i = 0; i < parts[p]->halo_dev.size(); i++) {
            const size_t ns = (size_t)parts[p]->halo_dev[i].n_send;
            void *bs = nullptr, *bv = nullptr;
            HIP_TRY(hipMalloc(&bs, std::max<size_t>(1, ns * parts[p]->halo_record(LFM_HALO_SOLVARS) * el)));
            g->allocs.push_back(bs);
            HIP_TRY(hipMalloc(&bv, std::max<size_t>(1, ns * parts[p]->halo_record(LFM_HALO_VISCOUS) * el)));
            g->allocs.push_back(bv);
            g->buf_s[p].push_back(bs);
            g->buf_v[p].push_back(bv);
        }
    }
    if (g->adjust) {
        std::vector<double *> mins(n), dts(n);
        for (int p = 0; p < n; p++) {
            mins[p] = parts[p]->d_dtmin;
            dts[p] = parts[p]->d_dt_ptr();
        }
        HIP_TRY(hipMalloc((void **)&g->d_dtmin, n * sizeof(double *)));
        g->allocs.push_back(g->d_dtmin);
        HIP_TRY(hipMalloc((void **)&g->d_dt, n * sizeof(double *)));
        g->allocs.push_back(g->d_dt);
        HIP_TRY(hipMemcpy(g->d_dtmin, mins.data(), n * sizeof(double *), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g->d_dt, dts.data(), n * sizeof(double *), hipMemcpyHostToDevice));
    }
    *out = g.release();
    return LFM_OK;
}

static int exchange(lfm_part_group *g, int kind) {
    const int n = (int)g->parts.size();
    auto &bufs = kind == LFM_HALO_SOLVARS ? g->buf_s : g->buf_v;
    for (int p = 0; p < n; p++)
        for (size_t i = 0; i < g->parts[p]->L.halo.size(); i++) {
            int rc = g->parts[p]->halo_pack((int)i, kind, bufs[p][i]);
            if (rc) return rc;
        }
    for (int q = 0; q < n; q++)
        for (size_t i = 0; i < g->parts[q]->L.halo.size(); i++) {
            const int r = g->parts[q]->L.halo[i].rank;
            int rc = g->parts[q]->halo_unpack((int)i, kind, bufs[r][g->peer[q][i]]);
            if (rc) return rc;
        }
    return LFM_OK;
}

int lfm_part_group_run(lfm_part_group *g, int nsteps) {
    if (!g || nsteps < 0) return fail(LFM_EINVAL, "bad argument");
    const int n = (int)g->parts.size();
    // every partition's earlier work is complete before the group stream
    // takes over its launches
    for (Solver *p : g->parts) HIP_TRY(hipStreamSynchronize(p->stream));
    std::vector<hipStream_t> own(n);
    for (int p = 0; p < n; p++) {
        own[p] = g->parts[p]->stream;
        g->parts[p]->stream = g->stream;
    }
    int rc = exchange(g, LFM_HALO_SOLVARS);  // halo ghosts hold the current state
    const int nst = g->parts[0]->nstages;
    for (int it = 0; it < nsteps && !rc; it++) {
        for (int rk = 0; rk < nst && !rc; rk++) {
            for (Solver *p : g->parts) p->stage_a(rk);
            rc = exchange(g, LFM_HALO_VISCOUS);
            if (rc) break;
            for (Solver *p : g->parts) p->stage_b(rk, true, !g->adjust);
            rc = exchange(g, LFM_HALO_SOLVARS);
        }
        if (rc) break;
        for (Solver *p : g->parts) p->end_step(!g->adjust);
        if (g->adjust) hipLaunchKernelGGL(lfm::group_dt_min, dim3(1), dim3(64), 0, g->stream, g->d_dtmin, g->d_dt, n);
        for (Solver *p : g->parts)
            if (p->post && !rc) rc = p->post_step();
    }
    hipError_t e = hipEventRecord(g->done, g->stream);
    for (int p = 0; p < n; p++) {
        g->parts[p]->stream = own[p];
        if (e == hipSuccess) e = hipStreamWaitEvent(own[p], g->done, 0);
    }
    if (rc) return rc;
    if (e != hipSuccess) return fail(LFM_EDEVICE, std::string("group stream hand-back: ") + hipGetErrorString(e));
    HIP_TRY(hipGetLastError());
    return LFM_OK;
}

int lfm_part_group_synchronize(lfm_part_group *g) {
    if (!g) return fail(LFM_EINVAL, "null handle");
    HIP_TRY(hipStreamSynchronize(g->stream));
    return LFM_OK;
}

int lfm_part_group_destroy(lfm_part_group *g) {
    delete g;
    return LFM_OK;
}

}  // extern "C"
