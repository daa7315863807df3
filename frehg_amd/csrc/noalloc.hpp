// Allocation-free stencil solve (frehg_ls_solve_noalloc): set-up and entry (noalloc.hip).
#pragma once

#include <cstddef>

#include "../../include/frehg_gpu.h"
#include "common.hpp"

namespace frehg {

// Size of the per-GPU scratch array the allocation-free solve works in (64 MiB: about 760k
// padded cells of a 3D operator).
constexpr size_t kScratchBytes = (size_t)64 << 20;

// Resolves the scratch array's device address on `device` (current on the calling thread); once
// per device, at context creation.
void noalloc_setup(int device);

// The solve itself; operator, rhs and solution are device arrays in the owned ordering. Throws
// ArgError / HipError / ConvergenceError like the other entry points' bodies.
void ls_solve_noalloc(int device, hipStream_t s, int ndim, int nx, int ny, int nz, const double* diag,
                      const double* const* offdiag, const double* b_owned, double* x_owned,
                      const frehg_solver_config& cfg, frehg_solve_stats* st);

}  // namespace frehg
