// scan.h -- device-wide exclusive scans over a value per index, for kernels
// whose sizes live in device memory: three grid-stride launches (per-block
// sums, a scan of the block sums, per-block scans with carries), enqueued
// without host round trips.  The functor F provides load() (reads the count
// from device memory), count() and operator()(i) -> u64 value.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gudni {
namespace devscan {
namespace {

constexpr int kScanBlocks = 512;    // grid of every scan launch
constexpr int kScanThreads = 256;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(v, (unsigned)o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan (blockDim.x a multiple of 64, at most 512).
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t &total)
{
    __shared__ uint64_t wave_sum[8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int waves = blockDim.x >> 6;
    const uint64_t incl = wave_incl_scan(v, lane);
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
    for (int k = 0; k < waves; ++k) {
        if (k < wave) pre += wave_sum[k];
        tot += wave_sum[k];
    }
    __syncthreads();
    total = tot;
    return pre + incl - v;
}

template <class F>
__global__ void __launch_bounds__(kScanThreads) scan_reduce(F f, uint64_t *partials)
{
    f.load();
    const int n = f.count();
    const int chunk = (n + kScanBlocks - 1) / kScanBlocks;
    const int b0 = min(n, (int)blockIdx.x * chunk), b1 = min(n, b0 + chunk);
    uint64_t s = 0;
    for (int i = b0 + (int)threadIdx.x; i < b1; i += kScanThreads) s += f(i);
    uint64_t total;
    (void)block_excl_scan(s, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanBlocks) scan_top(uint64_t *partials)
{
    const uint64_t v = partials[threadIdx.x];
    uint64_t total;
    const uint64_t ex = block_excl_scan(v, total);
    partials[threadIdx.x] = ex;
    if (threadIdx.x == 0) partials[kScanBlocks] = total;
}

template <class F>
__global__ void __launch_bounds__(kScanThreads) scan_down(F f, const uint64_t *partials, uint64_t *out)
{
    f.load();
    const int n = f.count();
    const int chunk = (n + kScanBlocks - 1) / kScanBlocks;
    const int b0 = min(n, (int)blockIdx.x * chunk), b1 = min(n, b0 + chunk);
    uint64_t carry = partials[blockIdx.x];
    for (int base = b0; base < b1; base += kScanThreads) {
        const int i = base + (int)threadIdx.x;
        const uint64_t v = i < b1 ? f(i) : 0;
        uint64_t total;
        const uint64_t ex = block_excl_scan(v, total);
        if (i < b1) out[i] = carry + ex;
        carry += total;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partials[kScanBlocks];
}

// The same scan by one block over the whole range (a kernel that handles a
// problem per block): out[i] for i < n and out[n] = the total.  Every thread
// of the block calls it.
template <class F>
__device__ void block_range_excl_scan(F f, uint64_t *out)
{
    f.load();
    const int n = f.count();
    uint64_t carry = 0;
    for (int base = 0; base < n; base += (int)blockDim.x) {
        const int i = base + (int)threadIdx.x;
        const uint64_t v = i < n ? f(i) : 0;
        uint64_t total;
        const uint64_t ex = block_excl_scan(v, total);
        if (i < n) out[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// out[i] = sum of f(j) for j < i, out[n] = the total; partials holds
// kScanBlocks + 1 words.
template <class F>
void excl_scan(const F &f, uint64_t *partials, uint64_t *out, hipStream_t stream)
{
    hipLaunchKernelGGL(scan_reduce<F>, dim3(kScanBlocks), dim3(kScanThreads), 0, stream, f, partials);
    hipLaunchKernelGGL(scan_top, dim3(1), dim3(kScanBlocks), 0, stream, partials);
    hipLaunchKernelGGL(scan_down<F>, dim3(kScanBlocks), dim3(kScanThreads), 0, stream, f,
                       (const uint64_t *)partials, out);
}

}  // namespace
}  // namespace devscan
}  // namespace gudni
