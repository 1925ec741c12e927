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
// essential columns in ascending rank (reduce_h1_gpu.py:495-502).  In a
// batch, class(birth) is a global edge index and each cloud's edges form one
// range, so one radix sort of every key orders the clouds as well; each
// row then goes to its cloud's block of the output.
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

// pre[s] = rows of the clouds before s; pre[nseg] = all rows.  One workgroup.
__global__ void k_seg_prefix(const int* __restrict__ counts, int nseg, int* __restrict__ pre) {
    __shared__ int wsum[4];
    const int t = threadIdx.x, lane = fph_lane(), wid = t / FPH_WAVE;
    int carry = 0;
    for (int b = 0; b <= nseg; b += 256) {
        const int i = b + t;
        const int v = i < nseg ? counts[i] : 0;
        int incl = v;
        for (int o = 1; o < FPH_WAVE; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == FPH_WAVE - 1) wsum[wid] = incl;
        __syncthreads();
        int before = carry, all = carry;
        for (int w = 0; w < 4; w++) {
            if (w < wid) before += wsum[w];
            all += wsum[w];
        }
        if (i <= nseg) pre[i] = before + incl - v;
        __syncthreads();
        carry = all;
    }
}

// Row k of the sorted keys.  One cloud: row k of the output.  A batch (pre
// not null): the cloud s is read off the key's birth class (an edge of
// cloud s, edges [seg_edges[s], seg_edges[s + 1])), and the row goes to
// seg_cols[s] + (k - pre[s]).
__global__ void k_diag_write(const unsigned long long* __restrict__ keys, long long K,
                             const int* __restrict__ counts, const int* __restrict__ pre,
                             const int* __restrict__ pivot,
                             const int* __restrict__ col_edge, long long row_base,
                             const int* __restrict__ row_edge, const float* __restrict__ edge_filt,
                             int ess_asc_all, const int* __restrict__ seg_cols,
                             const int* __restrict__ seg_edges,
                             const unsigned char* __restrict__ seg_ess_asc, int nseg,
                             float* __restrict__ out, int* __restrict__ out_edges) {
    const long long kk = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (kk >= K) return;
    const unsigned long long key = keys[kk];
    int s = 0;
    long long k = kk;
    if (pre) {
        if (kk >= pre[nseg]) return;
        s = segment_of(seg_edges, nseg, (long long)(key >> 33));
        k = seg_cols[s] + (kk - pre[s]);
    } else if (kk >= counts[0]) {
        return;
    }
    const int asc = seg_ess_asc ? seg_ess_asc[s] : ess_asc_all;
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
    if (nseg > 1 && !seg_edges) return FPH_ERR_ARG;
    const size_t k = (size_t)std::max<int64_t>(K, 1);
    // keys below E << 33, the unused ones all ones: the low 33 + bits(E)
    // bits order them
    int end_bit = 33;
    for (long long e = E; e > 0; e >>= 1) end_bit++;
    end_bit = std::min(end_bit, 64);
    size_t tmp = 0;
    FPH_CHECK_HIP(rocprim::radix_sort_keys(nullptr, tmp, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (unsigned)k, 0, end_bit,
                                           s));
    tmp = align_up(tmp);
    const size_t need = tmp + 2 * align_up(8 * k) + align_up(4 * ((size_t)nseg + 1));
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
    FPH_CHECK_HIP(rocprim::radix_sort_keys(ws, t, keys, sorted, (unsigned)K, 0, end_bit, s));
    int* pre = nullptr;
    if (seg_cols) {
        pre = (int*)(w + tmp + 2 * align_up(8 * k));
        k_seg_prefix<<<1, 256, 0, s>>>(counts, (int)nseg, pre);
        FPH_LAUNCH_CHECK();
    }
    k_diag_write<<<fph_blocks(K, 256), 256, 0, s>>>(sorted, K, counts, pre, pivot, col_edge,
                                                   row_base, row_edge, edge_filt, ess_ascending,
                                                   seg_cols, seg_edges, seg_ess_asc, (int)nseg,
                                                   out, out_edges);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
