// kernels_fp32p.hpp -- FP32 deviation-from-freestream time step (declarations).
#pragma once

#include "solver_state.hpp"

namespace lfm {

struct KArgs32 {
    int dummy;
};

template <int D>
void launch_bc32(Solver *) {}
template <int D>
void launch_gradvis32(Solver *) {}
template <int D>
void launch_flux32(Solver *) {}
template <int D>
void launch_update32(Solver *, double, double, bool, bool, bool) {}
template <int D>
void launch_cfl_dt32(Solver *, int, double, double *, double *) {}

}  // namespace lfm
