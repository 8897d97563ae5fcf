// batch.hip -- shared launches over several resident FP32-perturbation
// solvers of one scheme class, one RK step captured in a hipGraph.
//
// The reference has no serving mode; this is the device half of the
// continuous-batching service: the host side admits and retires solvers and
// re-forms the batch at step boundaries.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels_fp32p.hpp"
#include "lfm_hip.h"
#include "solver_state.hpp"

namespace lfm {

struct Batch {
    hipStream_t stream = nullptr;
    std::vector<Solver *> members;
    KArgs32 *d_args = nullptr;
    int32_t *d_maps = nullptr;
    size_t maps_cap = 0, args_cap = 0;
    BatchMap map_cells{}, map_faces{};
    int nb_cells = 0, nb_faces = 0;
    int D = 2, solver = 0, minmod = 0, nstages = 0;
    bool visc = true;
    // inviscid M1 / M2 members with their second buffers and face slots:
    // the stages after the first `separate` ones are one launch each
    bool fused = false;
    int separate = 0;
    double A[5]{}, B[5]{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~Batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_args) (void)hipFree(d_args);
        if (d_maps) (void)hipFree(d_maps);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

template <int D, int SOLVER>
static void enqueue_fused(Batch *b, int rk, int flip) {
    const dim3 blk(256);
    if (rk == b->nstages - 1)
        hipLaunchKernelGGL((stage_fused32_b<D, SOLVER, 1>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                           b->map_cells, (float)b->A[rk], (float)b->B[rk], rk == 0 ? 1 : 0, flip);
    else
        hipLaunchKernelGGL((stage_fused32_b<D, SOLVER, -1>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                           b->map_cells, (float)b->A[rk], (float)b->B[rk], rk == 0 ? 1 : 0, flip);
}

template <int D, bool V>
static void enqueue_step(Batch *b) {
    const dim3 blk(256);
    for (int rk = 0; rk < b->nstages; rk++) {
        if (b->fused && rk >= b->separate) {
            // an even number of one-launch stages: each step ends in a.q
            const int flip = (rk - b->separate) & 1;
            b->solver == 0 ? enqueue_fused<D, 0>(b, rk, flip) : enqueue_fused<D, 1>(b, rk, flip);
            continue;
        }
        // ghosts are refreshed by the update kernel (members join with fresh ghosts);
        // without viscosity solvers 0/1 consume no gradient or tensor
        if (V || b->solver == 2) {
            if (b->solver == 2 && b->minmod)
                hipLaunchKernelGGL((grad_vis32_b<D, 2, V, true>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                                   b->map_cells);
            else if (b->solver == 2)
                hipLaunchKernelGGL((grad_vis32_b<D, 0, V, true>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                                   b->map_cells);
            else
                hipLaunchKernelGGL((grad_vis32_b<D, 0, V, false>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                                   b->map_cells);
        }
        switch (b->solver) {
        case 0: hipLaunchKernelGGL((face_flux32_b<D, 0, V>), dim3(b->nb_faces), blk, 0, b->stream, b->d_args, b->map_faces); break;
        case 1: hipLaunchKernelGGL((face_flux32_b<D, 1, V>), dim3(b->nb_faces), blk, 0, b->stream, b->d_args, b->map_faces); break;
        default: hipLaunchKernelGGL((face_flux32_b<D, 2, V>), dim3(b->nb_faces), blk, 0, b->stream, b->d_args, b->map_faces); break;
        }
        // the last stage's update also forms the step's compute_cfl partials
        if (rk == b->nstages - 1)
            hipLaunchKernelGGL((cell_update32_b<D, 1>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                               b->map_cells, (float)b->A[rk], (float)b->B[rk], rk == 0 ? 1 : 0);
        else
            hipLaunchKernelGGL((cell_update32_b<D, -1>), dim3(b->nb_cells), blk, 0, b->stream, b->d_args,
                               b->map_cells, (float)b->A[rk], (float)b->B[rk], rk == 0 ? 1 : 0);
    }
    hipLaunchKernelGGL(cfl_final32_b, dim3((unsigned)b->members.size()), blk, 0, b->stream, b->d_args);
}

static int batch_set_locked(Batch *b, lfm_solver *const *hs, int n);

static int batch_set(Batch *b, lfm_solver *const *hs, int n) {
    // the one-launch stages' buffers of eligible members, allocated (under
    // the set-up mutex, which ensure_alt32 takes itself) before the batch
    // takes it; a member without them keeps the batch on separate launches
    for (int i = 0; i < n; i++)
        if (hs && hs[i] && hs[i]->impl && hs[i]->impl->fp32) (void)hs[i]->impl->ensure_alt32();
    std::lock_guard<std::mutex> lock(device_setup_mutex());
    return batch_set_locked(b, hs, n);
}

static int batch_set_locked(Batch *b, lfm_solver *const *hs, int n) {
    if (b->exec) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    b->members.clear();
    b->fused = false;
    if (n == 0) return LFM_OK;
    bool fused = true;
    std::vector<KArgs32> args;
    std::vector<int32_t> req_c, req_f, first_c, first_f;
    for (int i = 0; i < n; i++) {
        if (!hs[i] || !hs[i]->impl) return fail(LFM_EINVAL, "null solver in batch");
        Solver *s = hs[i]->impl.get();
        if (!s->fp32) return fail(LFM_EINVAL, "batched launches need FP32-perturbation solvers");
        if (!s->L.halo.empty()) return fail(LFM_EINVAL, "a partition with processor patches cannot join a batch");
        if (s->prm.have_res || s->prm.adjust_dt)
            return fail(LFM_EINVAL, "batched solvers run constant dt without residual output");
        if (i == 0) {
            b->D = s->L.D;
            b->solver = s->prm.solver;
            b->minmod = s->prm.minmod;
            b->nstages = s->nstages;
            b->visc = s->visc;
            for (int k = 0; k < 5; k++) {
                b->A[k] = s->A[k];
                b->B[k] = s->B[k];
            }
        } else if (s->L.D != b->D || s->prm.solver != b->solver || s->prm.minmod != b->minmod ||
                   s->nstages != b->nstages || s->visc != b->visc) {
            return fail(LFM_EINVAL, "batched solvers must share dimension, solver, minmod, rkOrder and whether "
                                    "they are viscous");
        }
        for (Solver *o : b->members)
            if (o == s) return fail(LFM_EINVAL, "solver listed twice in a batch");
        if (s->ghosts_stale) {
            s->bc();
            s->ghosts_stale = false;
        }
        HIP_TRY(hipStreamSynchronize(s->stream));
        b->members.push_back(s);
        args.push_back(*s->k32);
        args.back().q_alt = s->q_alt32;
        fused = fused && s->fused_ok && s->q_alt32 && s->k32->sl_nb && !s->post_forces;
        auto add = [&](std::vector<int32_t> &req, std::vector<int32_t> &first, long count) {
            first.push_back((int32_t)req.size());
            const long nb = (count + 255) / 256;
            for (long k = 0; k < nb; k++) req.push_back(i);
        };
        add(req_c, first_c, s->L.nc);
        add(req_f, first_f, s->L.nf);
    }
    b->nb_cells = (int)req_c.size();
    b->nb_faces = (int)req_f.size();
    b->fused = fused && b->solver < 2 && !b->visc;
    b->separate = b->nstages % 2;  // leaves an even number of one-launch stages
    const size_t nmap = req_c.size() + req_f.size() + 2 * (size_t)n;
    if (nmap > b->maps_cap) {
        if (b->d_maps) HIP_TRY(hipFree(b->d_maps));
        HIP_TRY(hipMalloc((void **)&b->d_maps, nmap * sizeof(int32_t)));
        b->maps_cap = nmap;
    }
    if ((size_t)n > b->args_cap) {
        if (b->d_args) HIP_TRY(hipFree(b->d_args));
        HIP_TRY(hipMalloc((void **)&b->d_args, n * sizeof(KArgs32)));
        b->args_cap = n;
    }
    std::vector<int32_t> maps;
    maps.reserve(nmap);
    auto put = [&](const std::vector<int32_t> &v) {
        const size_t off = maps.size();
        maps.insert(maps.end(), v.begin(), v.end());
        return b->d_maps + off;
    };
    b->map_cells = {put(req_c), nullptr};
    b->map_cells.first = put(first_c);
    b->map_faces = {put(req_f), nullptr};
    b->map_faces.first = put(first_f);
    HIP_TRY(hipMemcpy(b->d_maps, maps.data(), maps.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_args, args.data(), n * sizeof(KArgs32), hipMemcpyHostToDevice));
    // capture one RK step; on failure leave no capture or sticky error behind
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    if (b->D == 2)
        b->visc ? enqueue_step<2, true>(b) : enqueue_step<2, false>(b);
    else
        b->visc ? enqueue_step<3, true>(b) : enqueue_step<3, false>(b);
    hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
        b->exec = nullptr;
        b->members.clear();
        (void)hipGetLastError();
        return fail(LFM_EDEVICE, std::string("step graph capture: ") + hipGetErrorString(e));
    }
    return LFM_OK;
}

}  // namespace lfm

struct lfm_batch {
    lfm::Batch impl;
};

extern "C" {

int lfm_batch_create(lfm_batch **out) {
    if (!out) return lfm::fail(LFM_EINVAL, "null argument");
    auto *b = new lfm_batch;
    if (hipStreamCreateWithFlags(&b->impl.stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return lfm::fail(LFM_EDEVICE, "hipStreamCreateWithFlags failed");
    }
    *out = b;
    return LFM_OK;
}

int lfm_batch_destroy(lfm_batch *b) {
    if (!b) return LFM_OK;
    if (b->impl.stream) (void)hipStreamSynchronize(b->impl.stream);
    delete b;
    return LFM_OK;
}

int lfm_batch_set(lfm_batch *b, lfm_solver *const *solvers, int n) {
    if (!b || n < 0 || (n > 0 && !solvers)) return lfm::fail(LFM_EINVAL, "bad argument");
    return lfm::batch_set(&b->impl, solvers, n);
}

int lfm_batch_run(lfm_batch *b, int nsteps) {
    if (!b || nsteps < 0) return lfm::fail(LFM_EINVAL, "bad argument");
    if (b->impl.members.empty() || !b->impl.exec) return nsteps == 0 ? LFM_OK : lfm::fail(LFM_ESTATE, "empty batch");
    for (int i = 0; i < nsteps; i++) {
        hipError_t e = hipGraphLaunch(b->impl.exec, b->impl.stream);
        if (e != hipSuccess) return lfm::fail(LFM_EDEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    }
    for (lfm::Solver *s : b->impl.members) s->steps_done += nsteps;
    return LFM_OK;
}

int lfm_batch_synchronize(lfm_batch *b) {
    if (!b) return lfm::fail(LFM_EINVAL, "null handle");
    hipError_t e = hipStreamSynchronize(b->impl.stream);
    if (e != hipSuccess) return lfm::fail(LFM_EDEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return LFM_OK;
}

}  // extern "C"
