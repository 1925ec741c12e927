// Shared device helpers for the flash-ph MI355X (gfx950) kernels.
#pragma once

#include <cstring>  // rocprim's texture_cache_iterator uses memset without it
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flashph.h"

#define FPH_WAVE 64

#define FPH_CHECK_HIP(expr)                                  \
    do {                                                     \
        hipError_t err_ = (expr);                            \
        if (err_ != hipSuccess) return FPH_ERR_HIP;          \
    } while (0)

#define FPH_LAUNCH_CHECK()                                   \
    do {                                                     \
        hipError_t err_ = hipGetLastError();                 \
        if (err_ != hipSuccess) return FPH_ERR_HIP;          \
    } while (0)

// Value of autotune knob `knob` in effect for (size, d) (defined in fph_tune.hip).
int fph_tune_value(int knob, int64_t size, int64_t d);

// Sets, once per process, the LDS limit of the one-workgroup Boruvka form
// (fph_mst.hip); call outside stream capture.
hipError_t fph_mst_prepare();

// Capacity forms of the edge sort + vertex CSR (fph_edges.hip) and of the H0
// diagram (fph_mst.hip) for graph capture: sizes are the host capacity, the
// live edge count is read on the device, and no memset node is recorded.
int fph_edges_csr_capacity(uint64_t* keys, uint64_t* vals, const uint64_t* count_dev,
                           int64_t cap, int64_t n, int32_t* edge_i, int32_t* edge_j,
                           float* edge_dsq, float* edge_filt, int32_t* adj_ptr, int32_t* adj_idx,
                           int32_t* adj_rank, float* adj_dsq, void* ws, size_t* ws_bytes,
                           hipStream_t s);
int fph_h0_diagram_capacity(const uint8_t* mst_mask, const float* edge_dsq, int64_t cap,
                            int64_t n, float* h0, void* ws, size_t* ws_bytes, hipStream_t s);

static inline unsigned fph_blocks(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// First index in sorted x[0..n) with x[idx] >= v.
template <typename T>
__device__ __forceinline__ int fph_lower_bound(const T* x, int n, T v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Rank of undirected edge (s, t) from the vertex CSR, -1 if absent.
__device__ __forceinline__ int fph_edge_rank(const int* adj_ptr, const int* adj_idx,
                                             const int* adj_rank, int s, int t) {
    int lo = adj_ptr[s], hi = adj_ptr[s + 1];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        int x = adj_idx[mid];
        if (x < t) lo = mid + 1;
        else if (x > t) hi = mid;
        else return adj_rank[mid];
    }
    return -1;
}

__device__ __forceinline__ int fph_lane() { return (int)(threadIdx.x & (FPH_WAVE - 1)); }

__device__ __forceinline__ unsigned long long fph_lanemask_lt() {
    int l = fph_lane();
    return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}

// Reserves a workgroup's output range with one atomic on a shared counter
// and returns the calling wave's start: the waves' totals (`total`, uniform
// per wave) go through LDS, thread 0 adds their sum to *count, and wave w
// starts after the totals of waves 0..w-1.  One append per wave serialised
// on the counter at about 12 ns each on MI355X; per workgroup it is a
// quarter of that.  Contains barriers: every thread of the workgroup calls it.
template <int THREADS>
__device__ __forceinline__ unsigned long long fph_wg_reserve(int total,
                                                             unsigned long long* count) {
    __shared__ int s_wt[THREADS / FPH_WAVE];
    __shared__ unsigned long long s_base;
    const int wid = (int)threadIdx.x / FPH_WAVE;
    if (fph_lane() == 0) s_wt[wid] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 0;
        for (int w = 0; w < THREADS / FPH_WAVE; w++) all += s_wt[w];
        s_base = all ? atomicAdd(count, (unsigned long long)all) : 0ull;
    }
    __syncthreads();
    unsigned long long pos = s_base;
    for (int w = 0; w < wid; w++) pos += (unsigned long long)s_wt[w];
    return pos;
}

// Correctly rounded FP32 square root: sqrt in FP64 then round to FP32.  FP64
// carries 53 >= 2*24 + 2 bits, so the double rounding is exact for sqrt.
__device__ __forceinline__ float fph_sqrt_rn(float x) {
    return (float)__dsqrt_rn((double)x);
}
