// mpm_noalloc.hip — the binning sort of caller-provided keys without any device
// allocation: mpm_sort_pairs_host_noalloc (include/mpm/mpm.h).
//
// mpm_sort_pairs_host allocates its device keys, values, their ping-pong copies
// and the radix scratch on every call, and creates a stream.  This variant takes all
// of that from `scratch`, a file-scope device array of a fixed size (room for
// MPM_SORT_NOALLOC_MAX_KEYS keys), and runs on the null stream, so that callers
// which cannot allocate device memory can sort.  Each GPU holds its own copy of the
// array until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress) and kept per device; host code passes only that address to
// kernels and copies.  The array keeps the last call's contents; the sort needs
// nothing zeroed (every count and offset it reads it has written first).  Calls are
// serialised by a host mutex; calls on one GPU from other processes use their own
// copies.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/mpm/mpm.h"
#include "mpm_solver_impl.hpp"
#include "mpm_sort.hpp"

namespace {

constexpr size_t NOALLOC_WORDS = 4 * (size_t)MPM_SORT_NOALLOC_MAX_KEYS + mpm::radix_scratch_words(MPM_SORT_NOALLOC_MAX_KEYS);

// keys, values, their ping-pong copies (n words each), then the radix scratch
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint32_t* g_scratch[MAX_DEVICES] = {};   // device address of `scratch` on each GPU

int fail(int code, const std::string& msg) { return mpm::detail::set_error(code, msg); }

}  // namespace

extern "C" int mpm_sort_pairs_host_noalloc(const uint32_t* keys, int64_t n, int key_bits, uint32_t* out_keys,
                                           uint32_t* out_perm)
{
    if (n < 0) return fail(MPM_E_INVALID, "n out of range");
    if (n > MPM_SORT_NOALLOC_MAX_KEYS)
        return fail(MPM_E_INVALID, "n = " + std::to_string(n) + " exceeds the allocation-free sort's maximum of " +
                                       std::to_string(MPM_SORT_NOALLOC_MAX_KEYS) + " keys");
    if (key_bits < 1 || key_bits > 32) return fail(MPM_E_INVALID, "key_bits must be in [1, 32]");
    if (n > 0 && !keys) return fail(MPM_E_INVALID, "keys is NULL");
    if (n == 0) return MPM_OK;
    if (key_bits < 32)
        for (int64_t i = 0; i < n; ++i)
            if (keys[i] >> key_bits) return fail(MPM_E_INVALID, "key exceeds key_bits");

    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(MPM_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(MPM_E_HIP, "device index out of range");
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<uint32_t*>(p);
    }
    const int nn = (int)n;
    uint32_t* k = g_scratch[dev];
    uint32_t *v = k + nn, *ka = k + 2 * (size_t)nn, *va = k + 3 * (size_t)nn;
    uint32_t* rs = k + 4 * (size_t)nn;
    hipStream_t st = nullptr;
    if ((e = hipMemcpyAsync(k, keys, (size_t)nn * 4, hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail(MPM_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    bool in_alt = false;
    if ((e = mpm::radix_sort_pairs(k, v, ka, va, nn, key_bits, true, rs, st, &in_alt)) != hipSuccess)
        return fail(MPM_E_HIP, std::string("radix sort: ") + hipGetErrorString(e));
    if (out_keys && (e = hipMemcpyAsync(out_keys, in_alt ? ka : k, (size_t)nn * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return fail(MPM_E_HIP, std::string("download: ") + hipGetErrorString(e));
    if (out_perm && (e = hipMemcpyAsync(out_perm, in_alt ? va : v, (size_t)nn * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return fail(MPM_E_HIP, std::string("download: ") + hipGetErrorString(e));
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(MPM_E_HIP, std::string("sync: ") + hipGetErrorString(e));
    return MPM_OK;
}
