// util.hip -- fill and device-wide exclusive scan.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/bendy2d.h"
#include "kernels.h"

namespace bd {

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t stream) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
}

// Plain streaming copy, 16 bytes per lane and iteration: the reference kernel
// for reading HBM byte counters and for the achievable copy bandwidth.
__global__ void k_copy_f4(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanThreads * kScanPerThread;  // 1024 words per block

// Inclusive prefix sum across the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one 1024-word tile; the tile total goes to sums[block].
__global__ void __launch_bounds__(kScanThreads) k_scan_tile(const uint32_t *in, uint32_t *out, size_t n,
                                                            uint32_t *sums) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        local += v[k];
    }
    uint32_t inc = wave_inclusive_scan(local);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0;
    for (int w = 0; w < wave; w++) wave_off += s_wave[w];
    uint32_t run = wave_off + inc - local;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (sums && threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = run;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t *out, size_t n, const uint32_t *offs) {
    size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t o = offs[blockIdx.x];
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
        if (base + k < n) out[base + k] += o;
}

static int scan_level(const uint32_t *in, uint32_t *out, size_t n, hipStream_t stream, ScanTemp &tmp, int lvl,
                      uint64_t &launches, std::string &err) {
    size_t blocks = (n + kScanTile - 1) / kScanTile;
    if (blocks <= 1) {
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanThreads), 0, stream, in, out, n, (uint32_t *)nullptr);
        launches++;
        return BD_OK;
    }
    if (lvl >= 4) {
        err = "scan too large";
        return BD_ERR_CAPACITY;
    }
    BD_HIP(tmp.level[lvl].alloc(blocks));
    hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)blocks), dim3(kScanThreads), 0, stream, in, out, n, tmp.level[lvl].p);
    launches++;
    int rc = scan_level(tmp.level[lvl].p, tmp.level[lvl].p, blocks, stream, tmp, lvl + 1, launches, err);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)blocks), dim3(kScanThreads), 0, stream, out, n, tmp.level[lvl].p);
    launches++;
    return BD_OK;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, hipStream_t stream, ScanTemp &tmp,
                       uint64_t &launches, std::string &err) {
    if (!n) return BD_OK;
    return scan_level(in, out, n, stream, tmp, 0, launches, err);
}

}  // namespace bd

extern "C" int32_t bd_ext_copy_calibration(uint64_t bytes, float *ms) {
    if (bytes < 16 || (bytes & 15)) return BD_ERR_ARG;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev < 1) return BD_ERR_NODEVICE;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int32_t rc = BD_OK;
    if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) rc = BD_ERR_NOMEM;
    if (rc == BD_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = BD_ERR_DEVICE;
    if (rc == BD_OK) {
        const size_t n = bytes / 16;
        bd::launch_fill_u32((uint32_t *)src, bytes / 4, 0x3f800000u, nullptr);
        // one float4 per lane (the grid-stride loop only matters past 2^28 bytes)
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)1 << 20);
        float t = 0.0f;
        if (hipEventRecord(e0, nullptr) != hipSuccess) rc = BD_ERR_DEVICE;
        if (rc == BD_OK) hipLaunchKernelGGL(bd::k_copy_f4, dim3(blocks), dim3(256), 0, nullptr,
                                            (const float4 *)src, (float4 *)dst, n);
        if (rc == BD_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = BD_ERR_DEVICE;
        if (ms) *ms = t;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return rc;
}
