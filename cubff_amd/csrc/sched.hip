// Learned scheduler (MLP cost model + length buckets).  Placeholder until the
// model kernels land: every entry point reports that it is unavailable.
#include <hip/hip_runtime.h>

#include "sched.h"

namespace cubff {

int sched_ensure(SchedState &, const BatchDev &, size_t) { return -1; }
void sched_free(SchedState &) {}
hipError_t sched_order_list(hipStream_t, SchedState &, const BatchDev &, const uint32_t *,
                            uint32_t *, uint64_t, size_t) {
  return hipErrorNotSupported;
}
hipError_t sched_record(hipStream_t, SchedState &, const BatchDev &) {
  return hipErrorNotSupported;
}

}  // namespace cubff
