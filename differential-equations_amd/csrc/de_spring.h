// Spring-chain (ragged, uneven-size) ensemble launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/de_api.h"
#include "de_lane.h"

namespace de {
namespace parity {

// Launch a spring-chain ensemble.  `b` holds device pointers; host_sizes is a
// host copy of b->sizes (the launcher buckets systems by size).  Returns 0 on
// success, 1 if unsupported, -1 on HIP error, -4 on allocation failure.
int launch_spring(const de_ode_batch* b, const int32_t* host_sizes, const OdeArgs& a, int sched,
                  hipStream_t s);
int64_t spring_workspace_bytes(const de_ode_batch* b);
// Autotuner decision for the size class with cap masses: 0 and (G, C, block),
// or -1 if that class has not been tuned in this process.
int spring_tuned(int cap, int* g, int* c, int* block);

}  // namespace parity
}  // namespace de
