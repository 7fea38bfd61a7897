// Device memory of the allocation-free entry point (msg_propagate_noalloc).
//
// `scratch` is one file-scope device array of kScratchBytes.  Every GPU the
// library runs on holds its own copy until the process ends.  Host code never uses
// the array's name as an address: at context set-up (msg_ctx_create) the address of
// each device's copy is looked up once with hipGetSymbolAddress, with that device
// current, and kept per device; the allocation-free entry point carves its buffers
// out of that address.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "msg_internal.hpp"

namespace msg {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t scratch_address(int device, void **ptr) {
    static std::mutex mu;
    static std::map<int, void *> addr;  // device -> address of its copy of `scratch`
    std::lock_guard<std::mutex> lock(mu);
    const auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return hipSuccess;
    }
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess) return e;
    if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[device] = p;
    *ptr = p;
    return hipSuccess;
}

}  // namespace msg
