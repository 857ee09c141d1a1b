// Solver scratch for the entry point that allocates nothing
// (bw_batch_step_noalloc).
//
// The tolerance-mode solver's scratch (the per-cloth solver states, the
// per-slice FP64 partial sums and the group sums of large cloths) lives in
// one statically sized device array instead of memory the batch allocated.
// Each GPU the library runs on holds its own copy of the array until the
// process ends; its address is looked up once per device with
// hipGetSymbolAddress and only that address is used afterwards (the array's
// name in host code is a host-side stand-in, not a device address).  Every
// call on a GPU shares that GPU's array and finds the previous call's
// contents there: calls that overlap on one GPU race.
#include <mutex>
#include <stdexcept>
#include <string>

#include "device.h"
#include "noalloc.h"

namespace bw
{

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace
{
constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
void *g_ptr[MAX_DEVICES] = {};

__global__ void k_fill_zero(unsigned long long *p, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}
} // namespace

void *noalloc_scratch()
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES)
        throw std::runtime_error("state: device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[dev])
    {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("HIP error hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

void noalloc_clear(void *p, size_t bytes, hipStream_t s)
{
    const size_t n = (bytes + 7) / 8;
    if (n > 0)
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           (unsigned long long *)p, n);
}

} // namespace bw
