/*
 * mcx_atomic_bench.hip -- measures the device's scattered-atomic rate, the
 * roof that bounds the fluence tally (one atomic per voxel exit).
 *
 * Every lane performs `nper` atomic adds at uniformly random indices of an
 * array of `nelem` elements.  Modes:
 *   0  FP64 add, agent scope, no return   (the default tally's instruction)
 *   1  FP32 add, agent scope, no return
 *   2  FP32 add, agent scope, returning   (the FP32 tallies' instruction)
 *   3  FP64 add, workgroup scope, no return, into one private copy of the
 *      array per XCD (chosen by the XCC_ID hardware register), so the add
 *      can complete in that XCD's L2
 * and coalesced forms of mode 0/1, where groups of lanes share a random
 * aligned block and each lane adds to its own element of it:
 *   4  FP64, groups of 8 lanes per 64-byte block
 *   5  FP64, the whole wave on 64 consecutive elements (512 bytes)
 *   6  FP32 no return, groups of 16 lanes per 64-byte block
 * The array is zeroed first and its sum checked afterwards (every add adds
 * exactly 1), so a lost update is reported.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ __forceinline__ uint32_t xcc_id() {
    /* s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4); masked to 8 XCDs */
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u;
}

__device__ __forceinline__ uint32_t xs32(uint32_t& s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

template <int MODE>
__global__ void __launch_bounds__(256) atomic_bench(void* buf, size_t nelem, uint32_t nper, float* sink, uint32_t* xcchist) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    /* lanes of one group draw the same block index from a shared stream */
    constexpr uint32_t G = MODE == 4 ? 8u : (MODE == 5 ? 64u : (MODE == 6 ? 16u : 1u));
    uint32_t s = (gid / G) * 2654435761u + 12345u;
    const size_t nblk = nelem / G;
    const uint32_t x = xcc_id();
    if (threadIdx.x == 0) atomicAdd(xcchist + x, 1u);
    float acc = 0.f;
    for (uint32_t k = 0; k < nper; k++) {
        const size_t i = (size_t)(((uint64_t)xs32(s) * (uint64_t)nblk) >> 32) * G + gid % G;
        if (MODE == 0 || MODE == 4 || MODE == 5) {
            unsafeAtomicAdd((double*)buf + i, 1.0);
        } else if (MODE == 1 || MODE == 6) {
            unsafeAtomicAdd((float*)buf + i, 1.0f);
        } else if (MODE == 2) {
            acc += unsafeAtomicAdd((float*)buf + i, 1.0f);
        } else {
            double* copy = (double*)buf + (size_t)x * nelem;
            __hip_atomic_fetch_add(copy + i, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    if (MODE == 2 && acc == -1.f) sink[0] = acc;
}

__global__ void sum_f64(const double* a, size_t n, double* out) {
    double t = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) t += a[i];
    atomicAdd(out, t);
}

__global__ void sum_f32(const float* a, size_t n, double* out) {
    double t = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) t += a[i];
    atomicAdd(out, t);
}

}  // namespace

/* Returns 0 on success.  out[0] = atomics per second, out[1] = kernel ms,
 * out[2] = (sum of the array) / (atomics issued), 1 when no add was lost,
 * out[3..10] = workgroups seen per XCC id. */
extern "C" int mcx_amd_atomic_bench(int mode, size_t nelem, unsigned int nper, unsigned int nblocks, double* out) {
    if (mode < 0 || mode > 6 || nelem < 64 || nper == 0 || nblocks == 0) return -1;
    const size_t esize = (mode == 1 || mode == 2 || mode == 6) ? 4 : 8;
    const size_t copies = (mode == 3) ? 8 : 1;
    void* buf = nullptr;
    float* sink = nullptr;
    uint32_t* hist = nullptr;
    double* dsum = nullptr;
    hipEvent_t a, b;
    int rc = -2;
    if (hipMalloc(&buf, esize * nelem * copies) != hipSuccess) return -2;
    if (hipMalloc(&sink, 4) != hipSuccess || hipMalloc(&hist, 32) != hipSuccess || hipMalloc(&dsum, 8) != hipSuccess) goto done;
    (void)hipMemset(buf, 0, esize * nelem * copies);
    (void)hipMemset(hist, 0, 32);
    (void)hipMemset(dsum, 0, 8);
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, 0);
    switch (mode) {
        case 0: hipLaunchKernelGGL(atomic_bench<0>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
        case 1: hipLaunchKernelGGL(atomic_bench<1>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
        case 2: hipLaunchKernelGGL(atomic_bench<2>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
        case 3: hipLaunchKernelGGL(atomic_bench<3>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
        case 4: hipLaunchKernelGGL(atomic_bench<4>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
        case 5: hipLaunchKernelGGL(atomic_bench<5>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
        default: hipLaunchKernelGGL(atomic_bench<6>, dim3(nblocks), dim3(256), 0, 0, buf, nelem, nper, sink, hist); break;
    }
    (void)hipEventRecord(b, 0);
    if (hipEventSynchronize(b) == hipSuccess) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        const double n = (double)nblocks * 256.0 * nper;
        if (esize == 8)
            hipLaunchKernelGGL(sum_f64, dim3(1024), dim3(256), 0, 0, (const double*)buf, nelem * copies, dsum);
        else
            hipLaunchKernelGGL(sum_f32, dim3(1024), dim3(256), 0, 0, (const float*)buf, nelem * copies, dsum);
        double total = 0.0;
        uint32_t h[8];
        (void)hipMemcpy(&total, dsum, 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(h, hist, 32, hipMemcpyDeviceToHost);
        out[0] = n / (ms * 1e-3);
        out[1] = ms;
        out[2] = total / n;
        for (int i = 0; i < 8; i++) out[3 + i] = h[i];
        rc = hipGetLastError() == hipSuccess ? 0 : -3;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
done:
    if (buf) (void)hipFree(buf);
    if (sink) (void)hipFree(sink);
    if (hist) (void)hipFree(hist);
    if (dsum) (void)hipFree(dsum);
    return rc;
}
