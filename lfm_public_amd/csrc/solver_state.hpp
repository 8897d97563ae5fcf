// solver_state.hpp -- the device-resident state behind one lfm_solver handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels_fp64.hpp"
#include "lfm_hip.h"
#include "mesh_builder.hpp"

namespace lfm {

// FP32-perturbation kernel arguments (defined in kernels_fp32p.hpp).
struct KArgs32;
// Post-processing state (averages, wall forces; defined in post.hip).
struct PostState;
void free_post(PostState *p);

int fail(int code, const std::string &msg);

// Returns the LFM_EDEVICE status (with the failing call and HIP's message
// recorded as the last error) from the enclosing function when a HIP call
// fails.
#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(LFM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Serialises device set-up (allocation, synchronous uploads) of new solvers
// with hipGraph capture on other threads: a synchronous copy issued while a
// capture is open would invalidate it.
std::mutex &device_setup_mutex();

struct Solver {
    lfm_params prm{};
    SolverLayout L;
    Gas gas{};
    double A[5]{}, B[5]{};
    int nstages = 0;
    double dt = 0.0;
    bool fp32 = false;
    bool visc = true;  // mu != 0: the step forms the viscous terms

    int device = 0;  // the HIP device current when the handle was created
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *h_scalar = nullptr;   // pinned host scratch (64 doubles)
    std::vector<void *> allocs;
    int64_t device_bytes = 0;

    // device scalars: [0] deltaT, [2..6] residual sums, [8] cfl, [9] dt min
    double *d_scalar = nullptr;
    double *d_res = nullptr, *d_cfl = nullptr, *d_dtmin = nullptr;
    double *d_dt_ptr() { return d_scalar; }

    KArgs k64{};
    KArgs32 *k32 = nullptr;       // owned via allocs-free host block

    // per-kernel-group timing with HIP events (lfm_profile_steps)
    // bc, grad_vis, face_flux, cell_update (and one-launch stages), cfl_dt
    // (final reductions), the last stage's update forming the CFL/dt extrema
    static constexpr int kProfGroups = 6;
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;       // pairs (start, stop)
    std::vector<int> prof_grp;
    size_t prof_n = 0;
    void prof_mark(int grp, bool start);

    // halo lists on the device, one entry per L.halo neighbour
    struct HaloDev {
        int32_t *send_cells = nullptr;  // [n_send] local cells, ascending
        int32_t *recv_ghost = nullptr;  // [n_recv] ghost slot (combined index nc+g)
        int32_t *recv_src = nullptr;    // [n_recv] record in the received buffer
        int n_send = 0, n_recv = 0;
    };
    std::vector<HaloDev> halo_dev;
    bool halo_ready() const {
        for (const HaloNbr &h : L.halo)
            if (!h.imported) return false;
        return true;
    }
    int halo_index(int rank) const {
        for (size_t i = 0; i < L.halo.size(); i++)
            if (L.halo[i].rank == rank) return (int)i;
        return -1;
    }
    int halo_record(int kind) const {
        return ((kind & 1) ? L.D + 2 : 0) + ((kind & 2) ? 2 * L.D * L.D + 2 * L.D : 0);
    }
    int halo_pack(int ni, int kind, void *buf);
    int halo_unpack(int ni, int kind, const void *buf);
    int halo_upload_geometry();  // d, w of the faces after an import

    // post-processing (lfm_post_configure); post_forces mirrors
    // post->cfg.have_forces for the stage code
    PostState *post = nullptr;
    bool post_forces = false;
    // copies the boundary ghosts the wall forces read (their state as the
    // current stage's boundary fill left it)
    void post_snapshot();
    // the Mesh::solve post-processing after one step (time, averages, forces)
    int post_step();

    bool ghosts_stale = true;  // ghost states not derived from the current q
    double pending_A = 1.0;
    int pending_rk = 0;
    int64_t steps_done = 0;
    bool res_valid = false;

    ~Solver();

    // the end-of-step reduction: compute_dt (mode 0, min into out and, when
    // dt_dst is set, the deltaT the next step uses) or compute_cfl (mode 1)
    struct CflJob {
        int mode;
        double param;
        double *out, *dt_dst;
    };
    CflJob cfl_job(bool own_dt);
    bool cfl_fused = false;  // the last update of this step formed it

    // stage pieces on the handle's stream
    void bc();
    void gradvis(bool full, bool fin = false);
    void flux();
    // cfl: the update also forms that reduction from the updated cells (it
    // reads only a cell's own state and its faces' areas); not with res
    void update(double A, double B, bool zero_dq, bool res, bool accumulate_res, const CflJob *cfl = nullptr);
    void cfl_dt(int mode, double param, double *out, double *dt_dst);
    // the final single-block reduction after an update that formed the
    // block extrema of a CFL job
    void cfl_final(const CflJob &job);
    void step();
    // pieces of step(): stage part A (ghost fill if stale, gradients and
    // viscous tensors), part B (fluxes, update), and the end-of-step CFL/dt
    // reduction (dt written to the own deltaT when own_dt).  stage_b of the
    // last stage forms that reduction in the update when fuse_cfl is set,
    // the stage computes no residual and the path is FP32 or 2D; end_step
    // then launches nothing.
    void stage_a(int rk);
    void stage_b(int rk, bool fuse_cfl = false, bool own_dt = true);
    // FP32 inviscid M1 / M2 without halos or wall forces (fused_ok): stages
    // from fused_from() on are one launch each, reading one state buffer and
    // writing the other; after an odd count the two buffers swap roles, so
    // k32->q always holds the state between steps.  The second buffer is
    // allocated by the first step that uses it (ensure_alt32), so solvers
    // that only step through a batch never hold it.
    bool fused_ok = false;
    float *q_alt32 = nullptr;
    bool ensure_alt32();
    // per-cell face slots of the FP32 kernels (k32->sl_*), on first use
    bool slots_failed = false;
    bool ensure_slots32();
    int fused_from() const;
    void stage_fused(int rk, bool from_alt);
    // the compute_cfl final reduction carried into the next step: when
    // defer_cfl is set (by a loop that steps again at once) and the next
    // step begins with a launch that can take it (carry_target), the last
    // stage leaves its per-block CFL maxima in the reduction scratch and
    // that launch reduces them in one extra block (no stage reads the CFL
    // at a fixed deltaT), which saves the single-block launch.  flush_cfl()
    // launches a carried reduction on its own; the loops call it when they
    // end.
    bool defer_cfl = false;
    bool cfl_pending = false;
    int carry_target() const;
    void final_or_carry(const CflJob &job);
    void flush_cfl();
    void end_step(bool own_dt);
    void zero_state_for_step();

    int get_field(int which, double *out);
    int set_field(int which, const double *in);
    int check_finite(int64_t *n_bad);
    // the same count with its per-block partials in the library's static
    // device scratch instead of an allocated counter
    int check_finite_noalloc(int64_t *n_bad);
};

int create_fp32p(Solver *s);
void free_k32(KArgs32 *k);
int get_field32(Solver *s, int which, double *out);
int nonfinite32(Solver *s, unsigned long long *d);
// FP64 conservation integrals sum_c V_c q_c into out_dev[0..D+1] (device)
int integrals(Solver *s, double *out_dev);
int integrals64(Solver *s, double *out_dev);

}  // namespace lfm

// opaque handle of the C ABI
struct lfm_solver {
    std::unique_ptr<lfm::Solver> impl;
};
