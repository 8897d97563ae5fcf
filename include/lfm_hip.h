/*
 * lfm_hip.h -- C ABI of the MI355X (gfx950) libFastMesh time step.
 *
 * The entry points replace, one for one, the virtual time-step methods of
 * the reference's ISolver (api/cfdv0_solver.h:18-105) as Mesh<P,D>::solve
 * (src/mesh_solver.cpp:474-740) drives them, plus the set-up that
 * Mesh::initialize / Mesh::initializeSolver perform (src/mesh_reader.cpp:25,
 * src/mesh_solver.cpp:32-125).  One handle holds both reference submeshes
 * (boundary cells, then interior cells) of one partition in device memory.
 *
 * Conventions: every function returns 0 on success or a negative LFM_E*
 * code and never throws; lfm_last_error() gives the message of the last
 * failure on the calling thread.  Plain pointers and sizes only.  Host
 * arrays are copied; the caller keeps ownership.  Field arrays are SoA
 * [component][cell] in the ORIGINAL polyMesh cell order.  A handle must be
 * used by one host thread at a time; its work is ordered on one HIP stream.
 */
#ifndef LFM_HIP_H
#define LFM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LFM_OK 0
#define LFM_EINVAL (-1)   /* bad argument or unsupported configuration      */
#define LFM_EDEVICE (-2)  /* HIP runtime error                              */
#define LFM_ENOMEM (-3)   /* device or host allocation failed / over budget */
#define LFM_ESTATE (-4)   /* call out of order                              */
#define LFM_ENONFINITE (-5) /* non-finite value detected in the state        */

/* Boundary roles (init_boundary_conditions, src/cfd_v0.cpp:973-1005). */
#define LFM_BC_UNSET 0    /* non-empty patch in no BC list: ghost q = 0     */
#define LFM_BC_WALL 1
#define LFM_BC_INLET 2
#define LFM_BC_OUTLET 3
#define LFM_BC_EMPTY 4
#define LFM_BC_CYCLIC 5
#define LFM_BC_PROCESSOR 6 /* processor patch of a decomposed case: patch_nbr
                              holds the neighbour partition's rank */

/* Halo payloads (MPI_EXCHANGE_SPLIT / _PACKED, api/mpi_env.h:34-38): per
 * cell record, in the reference's send-buffer order (cfd_v0.cpp:3303-3331). */
#define LFM_HALO_SOLVARS 1 /* q: D+2                                        */
#define LFM_HALO_VISCOUS 2 /* dudx (D*D), dTdx (D), tauMC (D*D), sigmaU (D) */
#define LFM_HALO_PACKED 3  /* SOLVARS then VISCOUS                          */

/* One partition of a decomposed case (decomposePar): its rank and, per
 * boundary face (index f - n_internal), the global face key that both sides
 * of a processor patch share (|faceProcAddressing|,
 * polyMeshReaderOF.cpp:251-260).  Only processor-patch faces are read. */
typedef struct {
    int32_t rank;
    const int32_t *face_key;
} lfm_partition;

/* Precision of the device state. */
#define LFM_PREC_FP64 64  /* plain FP64, the reference's double build      */
#define LFM_PREC_FP32P 32 /* FP32 deviations from the freestream state     */

/* Field selectors for lfm_get_field / lfm_set_field. */
#define LFM_FIELD_Q 0       /* D+2 conservatives rho, rhoU, rhoE           */
#define LFM_FIELD_DQ 1      /* D+2 delta_q                                 */
#define LFM_FIELD_DUDX 2    /* D*D velocity gradient, row-major [i][j]     */
#define LFM_FIELD_DTDX 3    /* D                                           */
#define LFM_FIELD_TAUMC 4   /* D*D                                         */
#define LFM_FIELD_SIGMAU 5  /* D                                           */
#define LFM_FIELD_RHOGRAD 6 /* D   (solver 2 with minmod only)             */
#define LFM_FIELD_PGRAD 7   /* D   (solver 2 with minmod only)             */
#define LFM_FIELD_UGRAD 8   /* D*D (solver 2 with minmod only)             */
#define LFM_FIELD_RES 9     /* D+2 per-cell residual of the last step      */
#define LFM_FIELD_PRIM 16   /* D+2: p, U (D comps), T                      */

/* OpenFOAM polyMesh in memory (constant/polyMesh/{points,faces,owner,
 * neighbour,boundary}); patch roles resolved by the caller or by
 * lfm_patch_role(). */
typedef struct {
    int32_t n_points;
    const double *points;        /* [n_points*3]                          */
    int32_t n_faces;
    const int32_t *face_off;     /* [n_faces+1] CSR into face_pts         */
    const int32_t *face_pts;
    const int32_t *owner;        /* [n_faces]                             */
    int32_t n_internal;
    const int32_t *neighbour;    /* [n_internal]                          */
    int32_t n_cells;
    int32_t n_patches;
    const int32_t *patch_start;
    const int32_t *patch_size;
    const int32_t *patch_role;   /* LFM_BC_*                              */
    const int32_t *patch_nbr;    /* cyclic partner patch or -1            */
    const int32_t *cell_submesh; /* optional cellSubmesh override or NULL */
} lfm_polymesh;

/* Solver inputs: CInputReader keys (src/inputReader.cpp:13-64). */
typedef struct {
    int32_t dim;        /* fvSchemes lfm/dimension: 2 or 3                 */
    int32_t solver;     /* fvSchemes lfm/solver: 0 M1, 1 M2, 2 M2AUSM      */
    int32_t rk_order;   /* fvSchemes lfm/rkOrder: 0..5                     */
    int32_t minmod;     /* fvSchemes lfm/minmodExists                      */
    int32_t adjust_dt;  /* controlDict adjustTimeStep                      */
    int32_t have_res;   /* controlDict lfm/post/haveResiduals              */
    double max_co;      /* controlDict maxCo                               */
    double dt;          /* controlDict deltaT                              */
    double mol_weight, cp, mu, pr;   /* thermophysicalProperties          */
    double p_inf, T_inf, U_inf[3];   /* spongeDict pinf, Tinf, Uinf_*     */
    double Ls, k_sponge, mach;       /* spongeDict Ls, k, M               */
    int32_t precision;  /* LFM_PREC_FP64 or LFM_PREC_FP32P                 */
    int32_t laminar;    /* turbulenceProperties simulationType laminar: 1 */
} lfm_params;

typedef struct {
    int32_t dim, n_cells, n_cells_sm0, n_faces, n_ghosts, max_faces_per_cell;
    int32_t precision, solver;
    int64_t device_bytes;   /* device memory held by the handle; an FP32P
                               handle adds its face slots (3D gradients, or
                               inviscid M1/M2) and second state buffer
                               (inviscid M1/M2) when it first uses them   */
    double dt, cfl;         /* current deltaT and last convective CFL      */
    int64_t steps_done;
} lfm_info;

typedef struct lfm_solver lfm_solver;

/* Patch role from OpenFOAM patch type and name (cfd_v0.cpp:983-1002). */
int lfm_patch_role(const char *type, const char *name);

/* Mesh<P,D>::initialize + initializeSolver + ISolver::initialize: geometry
 * (polyMeshReaderOF), submesh split (polyMeshReaderOF.cpp:501-552),
 * neighbour typing and S/d/w (mesh_reader.cpp:163-615, cfd_v0.cpp:760-861),
 * ghost cells (cfd_v0.cpp:346-431), face order (cfd_v0.cpp:278-339) and
 * initial state (cfd_v0.cpp:867-967), uploaded to the current device.
 * alpha may be NULL when Ls <= 0. */
int lfm_solver_create(const lfm_polymesh *mesh, const lfm_params *prm, const double *p, const double *T,
                      const double *U3, const double *alpha, lfm_solver **out);
/* The same for one partition of a decomposed case: processor patches
 * (LFM_BC_PROCESSOR) get one halo ghost per face, filled by
 * lfm_halo_unpack.  The handle cannot step until lfm_halo_import has run
 * for every neighbour (the processor faces' d and w need the neighbour's
 * cell centres). */
int lfm_solver_create_part(const lfm_polymesh *mesh, const lfm_partition *part, const lfm_params *prm,
                           const double *p, const double *T, const double *U3, const double *alpha,
                           lfm_solver **out);
int lfm_solver_destroy(lfm_solver *s);
const char *lfm_last_error(void);
int lfm_get_info(lfm_solver *s, lfm_info *info);

/* ISolver::prepare_for_timestep (cfd_v0.cpp:1326-1333). */
int lfm_prepare_for_timestep(lfm_solver *s);
/* ISolver::prepare_for_RKstep (cfd_v0.cpp:1339-1377). */
int lfm_prepare_for_rkstep(lfm_solver *s, int rk);
/* ISolver::set_boundary_conditions (cfd_v0.cpp:1010-1221). */
int lfm_set_boundary_conditions(lfm_solver *s);
/* ISolver::calc_gradients / calc_gradients_M2AUSM (cfd_v0.cpp:1384-1567). */
int lfm_calc_gradients(lfm_solver *s);
/* ISolver::calc_VIS (cfd_v0.cpp:1744-1860). */
int lfm_calc_vis(lfm_solver *s);
/* ISolver::one_rk_step_{M1,M2,M2AUSM} for both submeshes (cfd_v0.cpp:1897-2832):
 * faces, residual scatter, sponge, RK update.  res (D+2, may be NULL) receives
 * the running sums of squared cell residuals when rk == 0 and have_res. */
int lfm_one_rk_step(lfm_solver *s, int rk, double dt, double *res);
/* ISolver::compute_dt / compute_cfl (cfd_v0.cpp:2887-2942), FP64 reductions. */
int lfm_compute_dt(lfm_solver *s, double max_co, double *dt_out);
int lfm_compute_cfl(lfm_solver *s, double dt, double *cfl_out);
/* n iterations of the Mesh::solve loop body (mesh_solver.cpp:474-740):
 * prepare, all RK stages, then dt (adjustTimeStep) or CFL; no host sync.
 * FP32P inviscid M1 / M2 solvers without halos or wall forces run most
 * stages as one launch (face fluxes formed in the update, the state
 * double-buffered); the results are the separate launches' bit for bit and
 * the state is in the handle's buffer after every step.  Between two steps
 * of one call, an FP32P step's compute_cfl final reduction (fixed deltaT)
 * runs as one extra block of the next step's first launch; the CFL that
 * lfm_get_info reports after the call is the last step's either way. */
int lfm_run(lfm_solver *s, int nsteps);
/* Wait for all queued work of the handle. */
int lfm_synchronize(lfm_solver *s);
/* Time n steps (as lfm_run launches them, without post-processing) with
 * HIP events on the handle's stream (ms). */
int lfm_time_steps(lfm_solver *s, int nsteps, double *ms_out);

/* Time n steps launched kernel by kernel with HIP events around every
 * kernel group on the handle's stream: ms_per_group[6] = bc fill,
 * gradients+viscous, face fluxes, cell update (+ residual reduction; the
 * one-launch stages of lfm_run), CFL/dt (the per-cell launch where it has
 * its own, and every final single-block reduction), the last stage's
 * update (or one-launch stage) that also forms the CFL/dt block extrema;
 * ms_total = whole. */
int lfm_profile_steps(lfm_solver *s, int nsteps, double *ms_per_group, double *ms_total);

/* Fields in original cell order, SoA [comp][cell], FP64 on the host. */
int lfm_get_field(lfm_solver *s, int which, double *out);
int lfm_set_field(lfm_solver *s, int which, const double *in);
/* Sum of squared residuals of the last step's stage 0 (D+2), as printed by
 * Mesh::solve (mesh_solver.cpp:763-793, before the square root). */
int lfm_get_residual(lfm_solver *s, double *res);
/* FP64 conservation integrals sum_c V_c q_c (D+2 values: mass, momentum,
 * energy), for either precision path. */
int lfm_conservation(lfm_solver *s, double *out);
/* Number of non-finite conservatives in the state (0 = healthy). */
int lfm_check_finite(lfm_solver *s, int64_t *n_bad);
/* The same count without allocating anything (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Its per-block partial counts live in a static device array of the
 * library (65,537 x 8 bytes per device, held until the process ends), so
 * the handle may have at most 16,777,216 cells (LFM_EINVAL beyond, with a
 * message naming the limit).  Every call on one device shares that array:
 * calls that overlap on one device (other host threads, other streams)
 * race; calls ordered one after another need no clearing between them. */
int lfm_check_finite_noalloc(lfm_solver *s, int64_t *n_bad);

/* ---------------------------------------------------------------------
 * Halo exchange between partitions (ISolver::mpi_communication / mpi_wait,
 * cfd_v0.cpp:3284-3692; set-up messages of calculateVNeighbour,
 * mesh_reader.cpp:512-612, and assign_pointers, cfd_v0.cpp:616-700).
 * --------------------------------------------------------------------- */
/* Neighbour ranks, in order of first appearance over the boundary-submesh
 * cells (ranks may be NULL to query n only). */
int lfm_halo_neighbours(lfm_solver *s, int32_t *n, int32_t *ranks);
/* The set-up message for neighbour `rank`: per processor face toward it,
 * the face key, the local owner cell (boundary-submesh index) and the
 * displacement cell centre - face centre (3 doubles).  NULL arrays: size
 * only. */
int lfm_halo_export(lfm_solver *s, int32_t rank, int32_t *n_faces, int32_t *keys, int32_t *cells, double *disp);
/* Apply the message lfm_halo_export produced on partition `rank`. */
int lfm_halo_import(lfm_solver *s, int32_t rank, int32_t n_faces, const int32_t *keys, const int32_t *cells,
                    const double *disp);
/* Bytes of one payload to send to / receive from neighbour `rank`
 * (8-byte values for FP64 solvers, 4-byte deviations for FP32P). */
int lfm_halo_sizes(lfm_solver *s, int32_t rank, int kind, int64_t *send_bytes, int64_t *recv_bytes);
/* Pack the send list's records into the device buffer dev_buf, and unpack
 * a received buffer into the halo ghosts; both are ordered on the handle's
 * stream (synchronise the sender before the receiver reads). */
int lfm_halo_pack(lfm_solver *s, int32_t rank, int kind, void *dev_buf);
int lfm_halo_unpack(lfm_solver *s, int32_t rank, int kind, const void *dev_buf);

/* Several partitions of one case resident on this device, stepped as
 * Mesh::solve steps one rank each (no counterpart as an object: the
 * reference runs one MPI rank per partition): per stage, gradients and
 * viscous tensors of every partition, viscous halo exchange, fluxes and
 * update, solvars exchange; per step the dt minimum over partitions when
 * adjustTimeStep is set.  parts[i] must be the partition of rank i with
 * every neighbour imported; all of one precision and scheme.  While in a
 * group, a partition must not be stepped on its own. */
typedef struct lfm_part_group lfm_part_group;
int lfm_part_group_create(lfm_solver *const *parts, int n, lfm_part_group **out);
int lfm_part_group_run(lfm_part_group *g, int nsteps);
int lfm_part_group_synchronize(lfm_part_group *g);
int lfm_part_group_destroy(lfm_part_group *g);

/* ---------------------------------------------------------------------
 * Post-processing on the device: running pressure averages and wall forces
 * (ISolver::prePostProc / postProcAverage / updateAverageField /
 * postProcForces, cfd_v0.cpp:2966-2970, 3056-3250, driven as Mesh::solve
 * drives them, mesh_solver.cpp:253-271, 333-358, 423-429, 740-756, 827-838).
 * Once configured, lfm_run and lfm_part_group_run do after every step what
 * Mesh::solve does: time += deltaT (the new deltaT when adjusting),
 * time_step++, postProcAverage(time_step - average_start), and a force row
 * when time_step % forces_every == 0.
 * --------------------------------------------------------------------- */
typedef struct {
    int32_t have_average;    /* controlDict lfm/post/haveAverage                 */
    int32_t average_start;   /* indexStartAverage = (int)((tStartAverage - time)
                                / deltaT), mesh_solver.cpp:254                    */
    int32_t have_forces;     /* lfm/post/haveForces                              */
    int32_t forces_every;    /* lfm/post/saveForcesStep (>= 1)                   */
    int32_t forces_own_face; /* 0: each wall entry reads the cell slot its face
                                had before reorder_faces, in the arrays as they
                                are after it (the reference's indexing); 1: the
                                wall face's own S, d and ghost                   */
    double time;             /* current time; the clock of the recorded rows     */
} lfm_post;
/* (Re)start post-processing: accumulators zeroed, step counter and clock
 * reset; have_* = 0 for both turns it off. */
int lfm_post_configure(lfm_solver *s, const lfm_post *cfg);
/* ISolver::postProcAverage(n): for n > 0, pAVG += p and pRMS += (p - pAVG/n)^2
 * per cell (FP64 accumulators on either precision path). */
int lfm_post_average(lfm_solver *s, int32_t n);
/* updateAverageField(n): pAvg = pAVG/n, pRMS = sqrt(pRMS/n), original cell
 * order (n > 0). */
int lfm_post_average_fields(lfm_solver *s, int32_t n, double *pavg, double *prms);
/* Wall patches (m_nWallBCList: patches of type wall, in patch order). */
int lfm_post_walls(lfm_solver *s, int32_t *n, int32_t *patches);
/* ISolver::postProcForces(time): record one row now.  The boundary ghosts are
 * read as the last boundary fill left them (after a step, the start of its
 * last stage); a state set from outside is filled first. */
int lfm_post_forces(lfm_solver *s, double time);
/* Drain up to max_rows recorded rows into rows (NULL: *n_rows = rows
 * pending, nothing drained).  Row layout: time, then per wall patch Fp[D]
 * and Fv[D] of this partition (a decomposed case sums the partitions, as
 * postProcForces' MPI_Allreduce does).  Rows are kept in recording order. */
int lfm_post_forces_read(lfm_solver *s, int32_t max_rows, double *rows, int32_t *n_rows);

/* ---------------------------------------------------------------------
 * Shared launches over several resident solvers (no counterpart in the
 * reference, which runs one solve per process): the FP32-perturbation
 * solvers of one scheme class (dimension, solver, minmod, rkOrder; constant
 * dt, no residual output) are stepped by single kernel launches covering all
 * of them, one RK step captured as a hipGraph and replayed.  Each solver
 * computes exactly what lfm_run would compute for it alone.  While a solver
 * is in a batch, only the batch may step it.
 * --------------------------------------------------------------------- */
typedef struct lfm_batch lfm_batch;
int lfm_batch_create(lfm_batch **out);
int lfm_batch_destroy(lfm_batch *b);
/* Replace the member set (may be empty) and capture the step graph.
 * Inviscid M1 / M2 members get the one-launch stages' second state buffer
 * and face slots here if they lack them (their lfm_info.device_bytes grows);
 * where that memory cannot be had the batch runs separate launches. */
int lfm_batch_set(lfm_batch *b, lfm_solver *const *solvers, int n);
/* Queue n steps of every member on the batch's stream. */
int lfm_batch_run(lfm_batch *b, int nsteps);
int lfm_batch_synchronize(lfm_batch *b);

#ifdef __cplusplus
}
#endif
#endif
