/*
 * mcx_device.h -- host-side HIP helpers shared by the single-simulation and
 * the batched launch paths.
 */
#ifndef MCX_AMD_DEVICE_H
#define MCX_AMD_DEVICE_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <string>

#include "mcx_internal.h"
#include "mcx_kernel.h"

namespace mcx {

#define HIP_CHECK(x)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess)                                                                             \
            throw mcx::Error(-(int)e_, std::string("HIP error (") + hipGetErrorString(e_) + ") in " #x); \
    } while (0)

/* device allocation released on scope exit */
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t n) {
        bytes = n;
        if (n) HIP_CHECK(hipMalloc(&p, n));
    }
    template <typename T>
    T* as() const { return (T*)p; }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
    }
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

/* throws for Config features this engine does not implement */
void check_supported(const Config* cfg);
int media_words(unsigned int mediabyte);
size_t media_words_per_voxel(unsigned int mediabyte);
/* waves per SIMD for a kernel using `vgprs` architectural VGPRs */
int waves_per_simd(int vgprs);
/* launch constants of one simulation: grid, source, physics, boundary and
 * detector fields; buffers, work split and time window are left zero */
void fill_kparam(const Config* cfg, unsigned int maxgate, unsigned int is2d, KParam& kp);
/* whether work counters are requested (mcx_amd_set_counters) */
int counters_enabled();
/* hot-cell box of at most `cells` cells at the source (KParam::hot_*;
 * hot_off is left to the caller) */
void hot_box_geometry(const Config* cfg, uint32_t cells, KParam& kp);
/* LDS bytes per co-resident group when `groups` share one CU, at most cap */
size_t lds_share(int device, unsigned int groups, size_t cap);

}  // namespace mcx

#endif
