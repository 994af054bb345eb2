// The lexicographic (reference-order) FP64 solve behind fadesim_gs64_solve.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "fadesim.h"

namespace fadesim {

// Workspace bytes gslex_solve needs for an n x m crossbar.
size_t gslex_workspace_bytes(int n, int m);
// Enqueues the solve of d on stream s with the workspace ws (at least
// gslex_workspace_bytes); d is validated by the caller.
int gslex_solve(const fadesim_gs64_desc *d, void *ws, hipStream_t s);

} // namespace fadesim
