// noalloc.hip -- the fixed device workspace of rc_solve_noalloc: a file-scope
// __device__ array that every GPU the library runs on holds for the life of
// the process, so that entry point never allocates device memory.
//
// Host code must not pass the array's name to a kernel (in host code it is
// only a stand-in for the symbol): noalloc_workspace() resolves the device
// address once per device with hipGetSymbolAddress, with that device current,
// and every later use goes through that address.
#include <mutex>

#include "kernels.h"

namespace rc {

namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {nullptr};

__global__ void k_zero_slot(double *v, int n)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) v[n] = 0.0;
}

}  // namespace

hipError_t noalloc_workspace(int device, void **addr)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;     // the caller makes the device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_addr[device] = p;
    }
    *addr = g_addr[device];
    return hipSuccess;
}

void launch_zero_slot(double *v, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_slot, dim3(1), dim3(64), 0, s, v, n);
}

}  // namespace rc
