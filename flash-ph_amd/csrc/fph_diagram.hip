// Diagram epilogue for H1 and H2.
//
// Replaces the assembly of reduce_h1_gpu.py:495-502 and 574-604: the
// reference concatenates
//     [apparent pairs, ascending column rank]
//  ++ [reduced pairs, descending column rank]      (processing order)
//  ++ [essential columns, descending column rank]  (processing order)
// and then sorts the rows stably by birth value.  Here every column gets one
// 64-bit key
//     class(birth) << 33 | group << 31 | order-in-group
// where class(birth) is the first edge index carrying the same FP32 birth
// value (edge values are non-decreasing in rank), so one radix sort yields
// exactly that row order.  With no triangle at all the reference lists the
// essential columns in ascending rank (reduce_h1_gpu.py:495-502).
#include "fph_common.h"

#include <rocprim/rocprim.hpp>

namespace {

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

__device__ __forceinline__ unsigned order_of(int group, long long c, int ess_asc) {
    return (group == 0 || (group == 2 && ess_asc)) ? (unsigned)c : (unsigned)(0x7FFFFFFFLL - c);
}

// Segment of column c (a point cloud of a batch); the whole range when
// seg_cols is null.
__device__ __forceinline__ int segment_of(const int* seg_cols, int nseg, long long c) {
    if (!seg_cols) return 0;
    int lo = 0, hi = nseg;  // last s with seg_cols[s] <= c
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg_cols[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_diag_keys(long long K, const unsigned char* __restrict__ kind,
                            const int* __restrict__ off, const unsigned* __restrict__ data,
                            const int* __restrict__ pivot, const int* __restrict__ col_edge,
                            long long row_base, const int* __restrict__ row_edge,
                            const float* __restrict__ edge_filt, long long E, int ess_asc_all,
                            const int* __restrict__ seg_cols, const int* __restrict__ seg_edges,
                            const unsigned char* __restrict__ seg_ess_asc, int nseg,
                            unsigned long long* __restrict__ keys, int* __restrict__ counts) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int group = -1, s = 0;
    if (c < K) {
        const unsigned char kd = kind[c];
        const int be = col_edge ? col_edge[c] : (int)c;
        if (kd == 1) {
            group = 2;
        } else if (kd == 2 || kd == 3) {
            const int p = pivot[c];
            if (p < 0) {
                group = 2;
            } else {
                const float b = edge_filt[be];
                const float dv = edge_filt[row_edge[(long long)p - row_base]];
                if (dv > b) group = kd == 2 ? 0 : 1;
            }
        }
        if (group < 0) {
            keys[c] = ~0ull;
        } else {
            s = segment_of(seg_cols, nseg, c);
            const long long e0 = seg_edges ? seg_edges[s] : 0;
            const long long e1 = seg_edges ? seg_edges[s + 1] : E;
            const int asc = seg_ess_asc ? seg_ess_asc[s] : ess_asc_all;
            const float bval = edge_filt[be];
            const unsigned long long cls = (unsigned long long)(
                e0 + fph_lower_bound(edge_filt + e0, (int)(e1 - e0), bval));
            keys[c] = (cls << 33) | ((unsigned long long)group << 31) | order_of(group, c, asc);
        }
    }
    // rows per segment: one atomic per wavefront and segment (consecutive
    // columns share their segment), not one per row on a shared counter
    bool pend = group >= 0;
    while (true) {
        const unsigned long long m = __ballot(pend);
        if (!m) break;
        const int s0 = __shfl(s, __ffsll((long long)m) - 1);
        const unsigned long long mine = __ballot(pend && s == s0);
        if (pend && s == s0) {
            if ((mine & fph_lanemask_lt()) == 0) atomicAdd(&counts[s0], __popcll(mine));
            pend = false;
        }
    }
}

__global__ void k_diag_write(const unsigned long long* __restrict__ keys, long long K,
                             const int* __restrict__ counts, const int* __restrict__ pivot,
                             const int* __restrict__ col_edge, long long row_base,
                             const int* __restrict__ row_edge, const float* __restrict__ edge_filt,
                             int ess_asc_all, const int* __restrict__ seg_cols,
                             const unsigned char* __restrict__ seg_ess_asc, int nseg,
                             float* __restrict__ out, int* __restrict__ out_edges) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int s = segment_of(seg_cols, nseg, k);
    const long long k0 = seg_cols ? seg_cols[s] : 0;
    if (k - k0 >= counts[s]) return;
    const int asc = seg_ess_asc ? seg_ess_asc[s] : ess_asc_all;
    const unsigned long long key = keys[k];
    const int group = (int)((key >> 31) & 3);
    const unsigned ord = (unsigned)(key & 0x7FFFFFFFull);
    const long long c = (group == 0 || (group == 2 && asc)) ? (long long)ord : 0x7FFFFFFFLL - ord;
    const int be = col_edge ? col_edge[c] : (int)c;
    const int de = group == 2 ? -1 : row_edge[(long long)pivot[c] - row_base];
    out[2 * k] = edge_filt[be];
    out[2 * k + 1] = de < 0 ? __builtin_inff() : edge_filt[de];
    if (out_edges) {
        out_edges[2 * k] = be;
        out_edges[2 * k + 1] = de;
    }
}

__global__ void k_zero_counts(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

extern "C" int fph_diagram(int64_t K, const uint8_t* kind, const int32_t* col_off,
                           const uint32_t* col_data, const int32_t* pivot,
                           const int32_t* col_edge, int64_t row_base, const int32_t* row_edge,
                           const float* edge_filt, int64_t E, int ess_ascending,
                           const int32_t* seg_cols, const int32_t* seg_edges,
                           const uint8_t* seg_ess_asc, int64_t nseg, float* out,
                           int32_t* out_edges, int32_t* counts, void* ws, size_t* ws_bytes,
                           void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || !ws_bytes || !counts || nseg < 1) return FPH_ERR_ARG;
    if (nseg > 1 && !seg_cols) return FPH_ERR_ARG;
    if (K >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    const size_t k = (size_t)std::max<int64_t>(K, 1);
    size_t tmp = 0;
    if (seg_cols) {
        FPH_CHECK_HIP(rocprim::segmented_radix_sort_keys(
            nullptr, tmp, (const unsigned long long*)nullptr, (unsigned long long*)nullptr,
            (unsigned)k, (unsigned)nseg, (const int*)nullptr, (const int*)nullptr, 0, 64, s));
    } else {
        FPH_CHECK_HIP(rocprim::radix_sort_keys(nullptr, tmp, (const unsigned long long*)nullptr,
                                               (unsigned long long*)nullptr, (unsigned)k, 0, 64,
                                               s));
    }
    tmp = align_up(tmp);
    const size_t need = tmp + 2 * align_up(8 * k);
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    k_zero_counts<<<fph_blocks(nseg, 256), 256, 0, s>>>(counts, (int)nseg);
    FPH_LAUNCH_CHECK();
    if (K == 0) return FPH_OK;
    char* w = (char*)ws;
    unsigned long long* keys = (unsigned long long*)(w + tmp);
    unsigned long long* sorted = (unsigned long long*)(w + tmp + align_up(8 * k));
    k_diag_keys<<<fph_blocks(K, 256), 256, 0, s>>>(K, kind, col_off, col_data, pivot, col_edge,
                                                  row_base, row_edge, edge_filt, E,
                                                  ess_ascending, seg_cols, seg_edges,
                                                  seg_ess_asc, (int)nseg, keys, counts);
    FPH_LAUNCH_CHECK();
    size_t t = tmp;
    if (seg_cols) {
        FPH_CHECK_HIP(rocprim::segmented_radix_sort_keys(ws, t, keys, sorted, (unsigned)K,
                                                         (unsigned)nseg, seg_cols, seg_cols + 1,
                                                         0, 64, s));
    } else {
        FPH_CHECK_HIP(rocprim::radix_sort_keys(ws, t, keys, sorted, (unsigned)K, 0, 64, s));
    }
    k_diag_write<<<fph_blocks(K, 256), 256, 0, s>>>(sorted, K, counts, pivot, col_edge, row_base,
                                                   row_edge, edge_filt, ess_ascending, seg_cols,
                                                   seg_ess_asc, (int)nseg, out, out_edges);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
