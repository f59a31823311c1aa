// pp_noalloc.hip -- the fixed device scratch of pp_spatstat_run_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) run a job on a
// resident cloud with every temporary -- center flags and lists, tiles, device
// tables, counters, the k-th neighbour categories -- carved from one
// file-scope device array instead of the library's own allocations.  Each
// device that runs such a job holds its own copy of the array (kNoAllocBytes)
// until the process ends.  Host code never uses the array's name as an
// address: its device address is looked up once per device, with that device
// current, and kept.
#include <mutex>

#include "pp_launch.h"

namespace pp {

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

hipError_t noalloc_scratch(int device, void **ptr, size_t *bytes)
{
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess)
            return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device)
            (void)hipSetDevice(cur);
        if (e != hipSuccess)
            return e;
        if (!p)
            return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    *bytes = kNoAllocBytes;
    return hipSuccess;
}

} // namespace pp
