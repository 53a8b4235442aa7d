// Device scratch of the entry point that allocates nothing
// (cacg_bcbcg_device_noalloc, include/cacg.h).
//
// One file-scope device array per GPU, loaded with the code object and held
// until the process ends.  Its host-side name is only a stand-in for the
// runtime, never a device address: the device address of each GPU's copy is
// looked up once per device with hipGetSymbolAddress and handed out from
// here; the host code uses that pointer for everything (kernel arguments,
// copies, fills).
#include "kcommon.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[CACG_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void *g_scratch_ptr[MAX_DEVICES];

}  // namespace

extern "C" int cacg_noalloc_scratch(void **d_ptr, long *bytes)
{
    if (!d_ptr || !bytes)
        return (int)hipErrorInvalidValue;
    *d_ptr = nullptr;
    *bytes = 0;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return (int)e;
    if (dev < 0 || dev >= MAX_DEVICES)
        return (int)hipErrorInvalidDevice;
    void *p = __atomic_load_n(&g_scratch_ptr[dev], __ATOMIC_ACQUIRE);
    if (!p) {
        // first call on this device, with it current: its copy's address
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return (int)e;
        if (!p)
            return (int)hipErrorInvalidSymbol;
        __atomic_store_n(&g_scratch_ptr[dev], p, __ATOMIC_RELEASE);
    }
    *d_ptr = p;
    *bytes = CACG_NOALLOC_SCRATCH_BYTES;
    return 0;
}
