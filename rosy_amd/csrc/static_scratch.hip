// The allocation-free frame's working memory: one static device array, of which every GPU the
// library runs on holds its own copy until the process ends.  Host code never passes the array's
// name to a kernel or a copy (on the host it only stands in for the symbol): each device's address
// is looked up once with hipGetSymbolAddress, with that device current, and used from then on.
#include <hip/hip_runtime.h>

#include <mutex>

#include "context.h"

namespace {

static __device__ __align__(256) unsigned char scratch[rsy::kStaticScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

namespace rsy {

hipError_t static_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device) return hipErrorInvalidDevice;  // the caller makes `device` current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

void* static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace rsy
