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
// of faces whose N is a boundary (not a processor) ghost carry kGhostFaceFlag.
constexpr int32_t kGhostFaceFlag = 1 << 30;
constexpr int32_t kFaceMask = kGhostFaceFlag - 1;

// The reference's two submeshes of one partition, flattened into one index
// space: cells [0, nsm0) are submesh 0 (boundary), [nsm0, nc) submesh 1,
// boundary ghosts [nc, nc+ng).  Faces are listed in the order the
// reference's one_rk_step loops visit them (cell by cell, valid slots in
// reorder_faces order), each once, oriented from P (the visiting cell) to N.
// One processor-patch neighbour of a partition: the reference's MPI ghost
// bookkeeping (cfd_v0.cpp:432-540, 587-748) and neighbour-centre exchange
// (mesh_reader.cpp:390-612).  Halo ghosts are one per processor face.
struct HaloNbr {
    int rank = -1;
    std::vector<int32_t> ghost;       // ghost index (0-based, >= ng_bnd) per processor face
    std::vector<int32_t> face;        // layout face index per processor face
    std::vector<int32_t> key;         // global face key per processor face
    std::vector<int32_t> cell;        // local owner cell (combined index) per processor face
    std::vector<double> xf, xc;       // [3*n] face and owner-cell centres
    std::vector<int32_t> send_cells;  // local cells adjacent to this rank, ascending
    std::vector<int32_t> recv_src;    // after import: record in the neighbour's send list, per face
    bool imported = false;
};

struct SolverLayout {
    int D = 2;
    int nc = 0, nsm0 = 0, ng = 0, nf = 0, max_faces = 0;
    int ng_bnd = 0;                     // boundary ghosts [0, ng_bnd); halo ghosts [ng_bnd, ng)
    int rank = -1;                      // partition rank (-1: undecomposed)
    std::vector<HaloNbr> halo;
    std::vector<int32_t> orig;          // combined -> original cell id
    std::vector<int32_t> fP, fN;        // [nf]
    std::vector<double> S, d, w;        // S,d: [D][nf] SoA; w: [nf]
    std::vector<double> vinv, sigma;    // [nc]
    // Smagorinsky constant -2 (0.16 V^(1/3))^2 each cell's stored tauMC ends
    // with (empty when laminar): its own, or for a cell with no face of its
    // own to process, that of the cell whose face touched it last
    std::vector<double> smag;
    std::vector<double> centre, volume; // [nc*3], [nc] (combined order)
    // gather lists: per cell, faces touching it in processing order;
    // entry f >= 0: the cell is P of face f; entry ~f: the cell is N.
    std::vector<int32_t> g_off, g_ent;
    // slot lists: per cell, all valid faces in reorder_faces slot order
    // (CFL / dt loops, cfd_v0.cpp:2895, 2921), same signed encoding.
    std::vector<int32_t> s_off, s_ent;
    // boundary ghosts: interior cell (combined index) and BC role
    std::vector<int32_t> ghost_cell, ghost_role;
    // per boundary ghost: its patch, and the face postProcForces reads for
    // it (cfd_v0.cpp:3169-3233): boundaries[patch][i] keeps the slot the
    // face had before reorder_faces, the cell's arrays are indexed after it,
    // so the face read is whichever moved into that slot (ref), or the
    // ghost's own face (own).  nb: combined index of the neighbour read;
    // S, d: [D][ng_bnd] as the cell sees that face.
    std::vector<int32_t> ghost_patch;
    std::vector<int32_t> patch_role;  // [n_patches] LFM_BC_* of the polyMesh patches
    std::vector<int32_t> gf_ref_nb;
    std::vector<double> gf_ref_S, gf_ref_d, gf_own_S, gf_own_d;
    // initial conservatives [(D+2)][nc] combined order
    std::vector<double> q0;
};

// Builds the layout and the initial state; returns an empty string on
// success or an error message.
std::string build_layout(const lfm_polymesh &m, const lfm_partition *part, const lfm_params &prm,
                         const double *p, const double *T, const double *U3, const double *alpha,
                         SolverLayout &L);

// Applies a neighbour's set-up message to halo entry `ni`: d and the linear
// weight of its processor faces (mesh_reader.cpp:577-605, cfd_v0.cpp:
// 800-855) and the receive order (the neighbour's ascending send list).
std::string halo_import(SolverLayout &L, int ni, int n, const int32_t *key, const int32_t *cell,
                        const double *disp);

// Gas constants derived as set_init_conditions does (cfd_v0.cpp:880-902).
struct Gas {
    double R, R_inv, gamma, gm1, cp, mu, pr_inv;
    double p_inf, T_inf, rho_inf, E_inf, U_inf[3];
};
Gas derive_gas(const lfm_params &prm);

// Runge-Kutta 2N tables (cfd_v0.cpp:86-128); returns the stage count.
int rk_tables(int rk_order, double A[5], double B[5]);

}  // namespace lfm
