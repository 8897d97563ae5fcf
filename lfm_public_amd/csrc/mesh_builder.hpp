// mesh_builder.hpp -- host set-up of one partition: polyMesh -> SoA solver
// layout in the reference's cell and face processing order.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lfm_hip.h"

namespace lfm {

constexpr int kMaxFaces = 6;        // FACE_CNT in {3,4,5,6} (src/mesh_solver.cpp:56-74)
constexpr int kGhostSubmesh = 10000; // MAX_MPI_RANKS, api/mpi_env.h:4
// gather-list encoding: entry f >= 0 (cell is P), ~f (cell is N); P entries
// of faces whose N is a boundary ghost carry kGhostFaceFlag.
constexpr int32_t kGhostFaceFlag = 1 << 30;
constexpr int32_t kFaceMask = kGhostFaceFlag - 1;

// The reference's two submeshes of one partition, flattened into one index
// space: cells [0, nsm0) are submesh 0 (boundary), [nsm0, nc) submesh 1,
// boundary ghosts [nc, nc+ng).  Faces are listed in the order the
// reference's one_rk_step loops visit them (cell by cell, valid slots in
// reorder_faces order), each once, oriented from P (the visiting cell) to N.
struct SolverLayout {
    int D = 2;
    int nc = 0, nsm0 = 0, ng = 0, nf = 0, max_faces = 0;
    std::vector<int32_t> orig;          // combined -> original cell id
    std::vector<int32_t> fP, fN;        // [nf]
    std::vector<double> S, d, w;        // S,d: [D][nf] SoA; w: [nf]
    std::vector<double> vinv, sigma;    // [nc]
    std::vector<double> centre, volume; // [nc*3], [nc] (combined order)
    // gather lists: per cell, faces touching it in processing order;
    // entry f >= 0: the cell is P of face f; entry ~f: the cell is N.
    std::vector<int32_t> g_off, g_ent;
    // slot lists: per cell, all valid faces in reorder_faces slot order
    // (CFL / dt loops, cfd_v0.cpp:2895, 2921), same signed encoding.
    std::vector<int32_t> s_off, s_ent;
    // boundary ghosts: interior cell (combined index) and BC role
    std::vector<int32_t> ghost_cell, ghost_role;
    // initial conservatives [(D+2)][nc] combined order
    std::vector<double> q0;
};

// Builds the layout and the initial state; returns an empty string on
// success or an error message.
std::string build_layout(const lfm_polymesh &m, const lfm_params &prm, const double *p, const double *T,
                         const double *U3, const double *alpha, SolverLayout &L);

// Gas constants derived as set_init_conditions does (cfd_v0.cpp:880-902).
struct Gas {
    double R, R_inv, gamma, gm1, cp, mu, pr_inv;
    double p_inf, T_inf, rho_inf, E_inf, U_inf[3];
};
Gas derive_gas(const lfm_params &prm);

// Runge-Kutta 2N tables (cfd_v0.cpp:86-128); returns the stage count.
int rk_tables(int rk_order, double A[5], double B[5]);

}  // namespace lfm
