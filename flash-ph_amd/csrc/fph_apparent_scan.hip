// Ripser-style apparent-pair scan for H1 (the reference's unwired
// apparent_pair_scan, kernels/apparent_pair_kernel.py:101-190, 200-313).
//
// A non-forest edge e = (u, v) of rank r_e is apparent when some common
// neighbour w has rank(u, w) < r_e and rank(v, w) < r_e: the triangle
// (u, v, w) then has e as its largest edge, and among the triangles on e
// the one with the smallest such w is the youngest cofacet (colex3 grows
// with w for fixed u < v).  The output is the flag and that triangle's colex
// index C(c,3) + C(b,2) + a of its sorted vertices.  This predicate ignores
// the forest, so its set is contained in the pipeline's forest-cleared one
// (fph_apparent, reduce_h1_gpu.py:159-223).
//
// One wavefront per edge: lanes take the elements w of the shorter
// neighbour list in ascending order and binary-search the longer list.  The
// reference's Triton loop walks the merge of the two lists for at most
// 2 * md steps (md = the maximum degree rounded up to a power of two in
// [64, 2048]); the pair (w, w) is examined at step lu(w) + lv(w) - c(w)
// (elements below w in each list, less the common ones, which advance both
// at once), so with a cap the scan keeps only w examined before the cap --
// the reference's false negatives beyond degree 2048.  Without a cap
// (cap <= 0) it equals the reference's Numba scan.
#include "fph_common.h"

namespace {

constexpr int kScanWaves = 4;

__global__ __launch_bounds__(kScanWaves * FPH_WAVE) void k_apparent_scan(
    const int* __restrict__ ei, const int* __restrict__ ej, const int* __restrict__ ranks,
    long long K, const int* __restrict__ adj_ptr, const int* __restrict__ adj_idx,
    const int* __restrict__ adj_rank, long long cap, int* __restrict__ flag,
    long long* __restrict__ colex) {
    const long long k = (long long)blockIdx.x * kScanWaves + threadIdx.x / FPH_WAVE;
    if (k >= K) return;  // wave-uniform
    const int lane = fph_lane();
    const int re = ranks[k];
    int u = ei[re], v = ej[re];
    if (u > v) { const int x = u; u = v; v = x; }
    const int us = adj_ptr[u], ue = adj_ptr[u + 1];
    const int vs = adj_ptr[v], ve = adj_ptr[v + 1];
    // A = shorter list, B = longer; positions are relative to each start
    const bool a_is_u = (ue - us) <= (ve - vs);
    const int as = a_is_u ? us : vs, an = a_is_u ? ue - us : ve - vs;
    const int bs = a_is_u ? vs : us, bn = a_is_u ? ve - vs : ue - us;
    int found_w = -1;
    int common = 0;  // matches among A's elements before this chunk
    for (int base = 0; base < an; base += FPH_WAVE) {
        const int ia = base + lane;
        bool match = false, ok = false;
        int w = 0, pb = 0;
        if (ia < an) {
            w = adj_idx[as + ia];
            // lower bound of w in B
            int lo = 0, hi = bn;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (adj_idx[bs + mid] < w) lo = mid + 1;
                else hi = mid;
            }
            pb = lo;
            match = pb < bn && adj_idx[bs + pb] == w;
            if (match) ok = adj_rank[as + ia] < re && adj_rank[bs + pb] < re;
        }
        const unsigned long long mm = __ballot(match);
        // common elements strictly below this lane's w
        const int c_before = common + __popcll(mm & fph_lanemask_lt());
        if (ok && cap > 0) ok = (long long)ia + pb - c_before < cap;
        const unsigned long long okm = __ballot(ok);
        if (okm) {
            found_w = __shfl(w, __ffsll((long long)okm) - 1);
            break;
        }
        common += __popcll(mm);
    }
    if (lane == 0) {
        flag[k] = found_w >= 0 ? 1 : 0;
        long long cx = 0;
        if (found_w >= 0) {
            long long a = u, b = v, c = found_w;
            if (a > b) { const long long x = a; a = b; b = x; }
            if (b > c) { const long long x = b; b = c; c = x; }
            if (a > b) { const long long x = a; a = b; b = x; }
            cx = c * (c - 1) * (c - 2) / 6 + b * (b - 1) / 2 + a;
        }
        colex[k] = cx;
    }
}

}  // namespace

extern "C" int fph_apparent_scan(const int32_t* edge_i, const int32_t* edge_j,
                                 const int32_t* nonmst_ranks, int64_t K,
                                 const int32_t* adj_ptr, const int32_t* adj_idx,
                                 const int32_t* adj_rank, int64_t max_degree,
                                 int32_t* is_apparent, int64_t* pivot_colex, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || (K > 0 && (!edge_i || !edge_j || !nonmst_ranks || !adj_ptr || !adj_idx ||
                            !adj_rank || !is_apparent || !pivot_colex)))
        return FPH_ERR_ARG;
    if (K == 0) return FPH_OK;
    long long cap = 0;  // max_degree <= 0: no step cap
    if (max_degree > 0) {
        long long md = 64;
        while (md < max_degree && md < 2048) md <<= 1;
        cap = 2 * md;
    }
    k_apparent_scan<<<fph_blocks(K, kScanWaves), kScanWaves * FPH_WAVE, 0, s>>>(
        edge_i, edge_j, nonmst_ranks, K, adj_ptr, adj_idx, adj_rank, cap, is_apparent,
        (long long*)pivot_colex);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
