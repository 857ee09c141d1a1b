// Statically sized device scratch for bw_batch_step_noalloc (noalloc.hip).
#pragma once
#include <cstddef>

#include <hip/hip_runtime.h>

namespace bw
{

// bytes of solver scratch the entry point that allocates nothing provides on
// each device: solver states, two slots of per-slice partial sums and of
// group sums (a batch of up to ~52,000 slices of 64 rows, 3.3M vertices)
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(4) << 20;

// this device's address of the scratch array (looked up once per device)
void *noalloc_scratch();
// zero `bytes` at p on stream s with a fill kernel
void noalloc_clear(void *p, size_t bytes, hipStream_t s);

} // namespace bw
