// scan.hpp -- device-wide exclusive prefix sums in three launches (tile
// sums, one-workgroup scan of the tile sums, tile apply), for arrays whose
// length may live on the device.  Element i is read through a loader
// functor and its exclusive prefix handed to a writer functor, so the same
// kernels scan plain arrays and fields of descriptor records.
#pragma once

#include "swift_common.hpp"

namespace swift {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanThreads * kScanPer;

template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T *lds_waves, T *total) {
    // 256 threads = 4 waves: wave-inclusive scan, then the wave totals
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
    }
    if (lane == 63) lds_waves[w] = x;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        T t = lds_waves[i];
        before += i < w ? t : T(0);
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__device__ __forceinline__ int scan_len(const int32_t *n_dev, int n_host) {
    return n_dev ? *n_dev : n_host;
}

template <class T, class Load>
__global__ void __launch_bounds__(kScanThreads)
scan_tile_sums(Load ld, const int32_t *n_dev, int n_host, T *sums) {
    __shared__ T wv[kScanThreads / 64];
    const int n = scan_len(n_dev, n_host);
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    if ((int64_t)blockIdx.x * kScanTile >= n) return;  // uniform per block
    T s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j)
        if (i0 + j < n) s += ld((int)(i0 + j));
    T tot;
    block_excl_scan<T>(s, wv, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// exclusive scan of the tile sums by one 1024-thread workgroup; base0
// (optional, device) offsets every prefix; the grand total goes to
// total_a / total_b when given
template <class T>
__global__ void __launch_bounds__(1024)
scan_tile_prefix(T *sums, const int32_t *n_dev, int n_host, const T *base0, T *total_a,
                 T *total_b) {
    __shared__ T part[1024];
    const int n = scan_len(n_dev, n_host);
    const int ntiles = (n + kScanTile - 1) / kScanTile;
    T carry = base0 ? *base0 : T(0);
    for (int c0 = 0; c0 < ntiles; c0 += 1024) {
        const int t = c0 + (int)threadIdx.x;
        T v = t < ntiles ? sums[t] : T(0);
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            T o = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : T(0);
            __syncthreads();
            part[threadIdx.x] += o;
            __syncthreads();
        }
        if (t < ntiles) sums[t] = carry + part[threadIdx.x] - v;
        carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (total_a) *total_a = carry;
        if (total_b) *total_b = carry;
    }
}

template <class T, class Load, class Store>
__global__ void __launch_bounds__(kScanThreads)
scan_tile_apply(Load ld, Store st, const int32_t *n_dev, int n_host, const T *sums) {
    __shared__ T wv[kScanThreads / 64];
    const int n = scan_len(n_dev, n_host);
    const int64_t i0 = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    if ((int64_t)blockIdx.x * kScanTile >= n) return;
    T v[kScanPer], s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        v[j] = i0 + j < n ? ld((int)(i0 + j)) : T(0);
        s += v[j];
    }
    T tot;
    T run = sums[blockIdx.x] + block_excl_scan<T>(s, wv, &tot);
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        if (i0 + j < n) st((int)(i0 + j), run);
        run += v[j];
    }
}

// Short arrays (n_max <= kScanSmall): the whole scan in one 1024-thread
// workgroup, 16 elements a thread, so one launch instead of three (small
// serving batches are launch-bound).
constexpr int kScanSmallPer = 16;
constexpr int kScanSmall = 1024 * kScanSmallPer;

template <class T, class Load, class Store>
__global__ void __launch_bounds__(1024)
scan_small(Load ld, Store st, const int32_t *n_dev, int n_host, const T *base0, T *total_a,
           T *total_b) {
    __shared__ T wsum[16];
    const int n = min(scan_len(n_dev, n_host), n_host);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = (int)threadIdx.x * kScanSmallPer;
    T v[kScanSmallPer], s = 0;
#pragma unroll
    for (int j = 0; j < kScanSmallPer; ++j) {
        v[j] = i0 + j < n ? ld(i0 + j) : T(0);
        s += v[j];
    }
    T x = s;  // wave-inclusive scan of the thread sums, then the wave totals
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        T t = wsum[i];
        before += i < w ? t : T(0);
        all += t;
    }
    const T base = base0 ? *base0 : T(0);
    T run = base + before + x - s;
#pragma unroll
    for (int j = 0; j < kScanSmallPer; ++j) {
        if (i0 + j < n) st(i0 + j, run);
        run += v[j];
    }
    if (threadIdx.x == 0) {
        if (total_a) *total_a = base + all;
        if (total_b) *total_b = base + all;
    }
}

// n_max bounds the length (host-known); n_dev, when given, is the actual
// length on the device.  sums needs n_max / kScanTile + 1 elements.
template <class T, class Load, class Store>
hipError_t launch_scan(Load ld, Store st, const int32_t *n_dev, int n_max, T *sums, const T *base0,
                       T *total_a, T *total_b, hipStream_t s) {
    if (n_max <= kScanSmall) {
        hipLaunchKernelGGL((scan_small<T, Load, Store>), dim3(1), dim3(1024), 0, s, ld, st, n_dev,
                           n_max, base0, total_a, total_b);
        return hipGetLastError();
    }
    const int tiles = (n_max + kScanTile - 1) / kScanTile;
    if (tiles > 0)
        hipLaunchKernelGGL((scan_tile_sums<T, Load>), dim3(tiles), dim3(kScanThreads), 0, s, ld,
                           n_dev, n_max, sums);
    hipLaunchKernelGGL((scan_tile_prefix<T>), dim3(1), dim3(1024), 0, s, sums, n_dev, n_max, base0,
                       total_a, total_b);
    if (tiles > 0)
        hipLaunchKernelGGL((scan_tile_apply<T, Load, Store>), dim3(tiles), dim3(kScanThreads), 0, s,
                           ld, st, n_dev, n_max, sums);
    return hipGetLastError();
}

template <class T>
struct LoadArr {
    const T *p;
    __device__ T operator()(int i) const { return p[i]; }
};
template <class T>
struct StoreArr {
    T *p;
    __device__ void operator()(int i, T v) const { p[i] = v; }
};

}  // namespace swift
