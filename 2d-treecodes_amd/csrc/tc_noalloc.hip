// tc_noalloc.hip -- the static device scratch behind tc_potential_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get a potential solve
// whose sorted sources, tree, target order and counters all live in one
// file-scope device array.  Each GPU the library runs on holds its own copy
// of the array (TC_NOALLOC_SCRATCH_BYTES) until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device with hipGetSymbolAddress (at
// context creation, with that device current) and only that address is ever
// handed to kernels or copies.
#include "tc_common.h"

#include <mutex>

namespace {

constexpr size_t TC_NOALLOC_SCRATCH_BYTES = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[TC_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {nullptr};

}  // namespace

size_t tc_noalloc_scratch_bytes() { return TC_NOALLOC_SCRATCH_BYTES; }

// Device address of `scratch` on `device`, which must be the current device.
void *tc_noalloc_scratch_lookup(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        throw TcError(-1, "treecode: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            throw TcError(-2, "treecode: no device address for the no-allocation scratch");
        g_ptr[device] = p;
    }
    return g_ptr[device];
}
