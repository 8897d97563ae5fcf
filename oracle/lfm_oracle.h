/*
 * lfm_oracle.h -- CPU restatement of libFastMesh's ISolver time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (lfm_public_amd/liblfm_hip.so) never links, calls or
 * falls back to it.
 *
 * Input is a raw OpenFOAM-style polyMesh (points, faces, owner, neighbour,
 * patches) plus the initial p, T, U, alpha fields, so the oracle does its own
 * geometry, submesh split, neighbour typing and coefficient set-up
 * independently of the product's host code.
 */
#ifndef LFM_ORACLE_H
#define LFM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary roles, resolved from the patch type/name by the caller with the
 * rule of init_boundary_conditions (src/cfd_v0.cpp:973-1005). */
enum {
    LFMO_BC_UNSET   = 0, /* non-empty patch in no list: ghost q stays 0   */
    LFMO_BC_WALL    = 1,
    LFMO_BC_INLET   = 2,
    LFMO_BC_OUTLET  = 3,
    LFMO_BC_EMPTY   = 4, /* faces masked out (polyMeshReaderOF.cpp:47-59) */
    LFMO_BC_CYCLIC  = 5, /* coupled to patch_nbr (polyMeshReaderOF.cpp:226-248) */
    LFMO_BC_PROCESSOR = 6 /* processor patch: patch_nbr is the neighbour rank */
};

typedef struct {
    int n_points;
    const double *points;      /* [n_points*3]                      */
    int n_faces;
    const int *face_off;       /* [n_faces+1] CSR into face_pts     */
    const int *face_pts;
    const int *owner;          /* [n_faces]                         */
    int n_internal;
    const int *neighbour;      /* [n_internal]                      */
    int n_cells;
    int n_patches;
    const int *patch_start;    /* [n_patches]                       */
    const int *patch_size;
    const int *patch_role;     /* LFMO_BC_*                         */
    const int *patch_nbr;      /* cyclic partner patch or -1        */
    const int *cell_submesh;   /* optional override, NULL = rule    */
    const int *face_key;       /* per boundary face: global face id  */
                               /* (processor faces), NULL if serial  */
    int rank;                  /* this partition's rank (-1 serial)  */
} lfmo_mesh;

typedef struct {
    int dim;           /* 2 or 3                                   */
    int solver;        /* 0 = M1, 1 = M2, 2 = M2AUSM               */
    int rk_order;      /* 0..5                                     */
    int minmod;        /* minmodExists                             */
    int adjust_dt;     /* controlDict adjustTimeStep               */
    int have_res;      /* lfm/post haveResiduals                   */
    double max_co;
    double dt;         /* controlDict deltaT                       */
    double mol_weight, cp, mu, pr;
    double p_inf, T_inf, U_inf[3];
    double Ls, k_sponge, mach;
    int laminar;       /* turbulenceProperties simulationType laminar;
                          0 selects calc_VIS_Smagorinsky            */
} lfmo_params;

/* Field selectors for lfmo_get_field (all in original cell order). */
enum {
    LFMO_F_Q = 0,       /* (D+2) comps                                  */
    LFMO_F_DQ = 1,      /* (D+2)                                        */
    LFMO_F_DUDX = 2,    /* D*D, row-major dudx[i][j]                    */
    LFMO_F_DTDX = 3,    /* D                                            */
    LFMO_F_TAUMC = 4,   /* D*D                                          */
    LFMO_F_SIGMAU = 5,  /* D                                            */
    LFMO_F_RHOGRAD = 6, /* D                                            */
    LFMO_F_PGRAD = 7,   /* D                                            */
    LFMO_F_UGRAD = 8,   /* D*D                                          */
    LFMO_F_RES = 9,     /* (D+2)                                        */
    LFMO_F_VOLINV = 10, /* 1                                            */
    LFMO_F_SIGMA = 11,  /* 1  sponge sigma                              */
    LFMO_F_CENTRE = 12, /* 3  cell centre                               */
    LFMO_F_VOLUME = 13, /* 1                                            */
    LFMO_F_SUBMESH = 14, /* 1                                           */
    /* gradients calc_gradients[_M2AUSM] forms that no flux consumes
     * (cfd_v0.cpp:1384-1567); read by the oracle's own pins only */
    LFMO_F_RHOUGRAD = 15, /* D*D, row-major [j][i] = d(rho U_j)/dx_i      */
    LFMO_F_RHOEGRAD = 16, /* D                                          */
    LFMO_F_RPSIGRAD = 17, /* D                                          */
    LFMO_F_CGRAD = 18     /* D                                          */
};

void  *lfmo_create(const lfmo_mesh *mesh, const lfmo_params *prm,
                   const double *p, const double *T, const double *U3,
                   const double *alpha, char *err, int errlen);
void   lfmo_destroy(void *h);
/* Advance n time steps (Mesh::solve loop body, src/mesh_solver.cpp:474-740). */
int    lfmo_run(void *h, int nsteps);
/* One stage only (for per-stage parity); begin_step performs prepare_for_timestep. */
int    lfmo_stage(void *h, int rk, int begin_step);
int    lfmo_get_field(void *h, int which, double *out);
/* Overwrite the cells' q or dq (LFMO_F_Q / LFMO_F_DQ), original cell order;
 * ghosts follow at the next boundary fill. */
int    lfmo_set_field(void *h, int which, const double *in);
int    lfmo_ncomp(void *h, int which);
double lfmo_get_dt(void *h);
double lfmo_get_cfl(void *h);
int    lfmo_get_res(void *h, double *res);   /* D+2 sums of squares (last step) */
double lfmo_compute_dt(void *h, double max_co);
double lfmo_compute_cfl(void *h, double dt);
/* Per (cell, slot) geometry after reorder_faces: counts, neighbour original
 * cell id (or -1-ghost), S, d, w.  Arrays sized n_cells*6(*3). */
int    lfmo_get_slots(void *h, int *nface, int *nvalid, int *nbr,
                      double *S, double *d, double *w);
int    lfmo_n_ghost(void *h);
/* Boundary ghosts as set_boundary_conditions last left them: q
 * [(D+2)][ng], the original id of the interior cell each belongs to, and
 * its patch role (LFMO_BC_*).  Any array may be NULL. */
int    lfmo_get_ghosts(void *h, double *q, int *cell, int *role);

/* Partitions (src/cfd_v0.cpp:432-748 MPI ghosts and send lists;
 * src/mesh_reader.cpp:390-612 neighbour centres; src/cfd_v0.cpp:3284-3692
 * payloads).  Halo ghosts are one per neighbour cell, as ghost_mpi. */
int    lfmo_halo_nbrs(void *h, int *ranks);                 /* returns n      */
int    lfmo_halo_export(void *h, int rank, int *keys, int *cells, double *disp); /* returns n */
int    lfmo_halo_import(void *h, int rank, int n, const int *keys, const int *cells, const double *disp);
/* kind: 1 solvars, 2 viscous, 3 packed; bit 4 adds the M2AUSM minmod
 * gradients (rho, p, U), which no reference payload carries; returns the
 * number of doubles */
int    lfmo_halo_pack(void *h, int rank, int kind, double *buf);
int    lfmo_halo_unpack(void *h, int rank, int kind, const double *buf);
/* Stage halves: phase 0 = prepare (+prepare_for_timestep when begin_step),
 * boundary conditions, gradients, calc_VIS; phase 1 = one_rk_step. */
int    lfmo_stage_phase(void *h, int rk, int begin_step, int phase);
void   lfmo_set_dt(void *h, double dt);

/* Post-processing (cfd_v0.cpp:2966-2970, 3083-3103, 3169-3233). */
int    lfmo_set_boundary_conditions(void *h);
int    lfmo_average(void *h, int n);                              /* postProcAverage(n)  */
int    lfmo_average_fields(void *h, int n, double *pavg, double *prms);
int    lfmo_walls(void *h, int *patches);                         /* returns the count   */
/* per wall patch: Fp[D] then Fv[D]; own_face 0 = the reference's slot lookup */
int    lfmo_forces(void *h, int own_face, double *out);

/* Unit-level pieces (closed-form tests of the oracle itself). */
double lfmo_minmod(int D, double cphi, double aphi, const double *grad, const double *d, double wl, double flux);
void   lfmo_p5(double M, double alpha, double *pos_neg);
int    lfmo_face_rhs(const lfmo_params *prm, const double *qP, const double *qN, const double *S, const double *d,
                     double w, int ghost, const double *visP, const double *visN, double *rhs);

#ifdef __cplusplus
}
#endif
#endif
