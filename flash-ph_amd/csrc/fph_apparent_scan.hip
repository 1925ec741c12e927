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

#include <climits>

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

// ---------------------------------------------------------------- H2
//
// The reference's unwired h2_apparent_pair_scan (kernels/
// apparent_pair_h2_kernel.py:161-317, 328-441).  Triangle t = (a, b, c),
// a < b < c, whose largest edge has rank r_t; its cofacets are the
// tetrahedra t + {w} for common neighbours w of a, b and c, ordered by
// (largest edge rank, colex4).  Every cofacet's largest rank is >= r_t, and
// equals r_t iff the ranks of (a, w), (b, w), (c, w) are all below r_t (no
// other edge has rank r_t).  colex4 of t + {w} grows with w, so the
// youngest cofacet has rank r_t iff such a w exists, and is then t + {w}
// for the smallest one.  t is apparent when, in addition, t is the oldest
// face of that tetrahedron: of its two faces on the largest edge (x, y),
// t = (x, y, s) and (x, y, w), t has the larger colex3, i.e. w < s.
//
// The reference's GPU loop walks the three-way merge of the neighbour
// lists for at most 3 md steps (md as in the H1 scan); each step advances
// every pointer below the current maximum, or all three on a match.  The
// cap can only cut a walk when la + lb + lc > 3 md; such triangles are
// walked by one lane, a maximum at a time (the lagging pointers' next
// positions by binary search), with the reference's step count.  All other
// triangles: lanes take the elements of the shortest list in ascending
// order and binary-search the other two.

__device__ __forceinline__ int lower_bound_from(const int* __restrict__ idx, int s, int p, int e,
                                                int w) {
    int lo = s + p, hi = e;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (idx[mid] < w) lo = mid + 1;
        else hi = mid;
    }
    return lo - s;
}

// rank of edge (x, y) from x's sorted neighbour list; -1 if absent
__device__ int edge_rank(const int* __restrict__ adj_ptr, const int* __restrict__ adj_idx,
                         const int* __restrict__ adj_rank, int x, int y) {
    const int s = adj_ptr[x], e = adj_ptr[x + 1];
    const int p = s + lower_bound_from(adj_idx, s, 0, e, y);
    return p < e && adj_idx[p] == y ? adj_rank[p] : -1;
}

// The reference loop with its step cap, one maximum at a time.  Returns the
// first common neighbour examined within the cap whose three ranks are
// below rt, or -1.
__device__ int walk_capped(const int* __restrict__ adj_idx, const int* __restrict__ adj_rank,
                           const int* st, const int* len, int rt, long long cap) {
    int p[3] = {0, 0, 0};
    long long step = 0;
    while (step < cap && p[0] < len[0] && p[1] < len[1] && p[2] < len[2]) {
        int v[3];
        for (int q = 0; q < 3; q++) v[q] = adj_idx[st[q] + p[q]];
        if (v[0] == v[1] && v[1] == v[2]) {  // this step examines the match
            if (adj_rank[st[0] + p[0]] < rt && adj_rank[st[1] + p[1]] < rt &&
                adj_rank[st[2] + p[2]] < rt)
                return v[0];
            p[0]++; p[1]++; p[2]++;
            step++;
            continue;
        }
        const int m = max(v[0], max(v[1], v[2]));
        // lagging pointers advance once per step until they reach a value
        // >= m; the maximum changes when one passes m or runs out
        int to[3];
        long long event = LLONG_MAX, all_at_m = 0;
        for (int q = 0; q < 3; q++) {
            to[q] = p[q];
            if (v[q] >= m) continue;
            to[q] = lower_bound_from(adj_idx, st[q], p[q], st[q] + len[q], m);
            const long long dq = to[q] - p[q];
            if (to[q] == len[q] || adj_idx[st[q] + to[q]] > m) event = min(event, dq);
            else all_at_m = max(all_at_m, dq);
        }
        if (event == LLONG_MAX) {  // every pointer reaches m: a match
            for (int q = 0; q < 3; q++) p[q] = to[q];
            step += all_at_m;
            continue;
        }
        for (int q = 0; q < 3; q++)
            if (v[q] < m) p[q] += (int)min((long long)(to[q] - p[q]), event);
        step += event;
    }
    return -1;
}

constexpr int kScanWavesH2 = 4;

__global__ __launch_bounds__(kScanWavesH2 * FPH_WAVE) void k_apparent_scan_h2(
    const int* __restrict__ tv0, const int* __restrict__ tv1, const int* __restrict__ tv2,
    const int* __restrict__ tmax, long long T, const int* __restrict__ adj_ptr,
    const int* __restrict__ adj_idx, const int* __restrict__ adj_rank, long long cap,
    int* __restrict__ flag, long long* __restrict__ colex) {
    const long long t = (long long)blockIdx.x * kScanWavesH2 + threadIdx.x / FPH_WAVE;
    if (t >= T) return;  // wave-uniform
    const int lane = fph_lane();
    const int a = tv0[t], b = tv1[t], c = tv2[t], rt = tmax[t];
    int st[3] = {adj_ptr[a], adj_ptr[b], adj_ptr[c]};
    int len[3] = {adj_ptr[a + 1] - st[0], adj_ptr[b + 1] - st[1], adj_ptr[c + 1] - st[2]};
    int found = -1;
    if (cap > 0 && (long long)len[0] + len[1] + len[2] > cap) {
        if (lane == 0) found = walk_capped(adj_idx, adj_rank, st, len, rt, cap);
        found = __shfl(found, 0);
    } else {
        // A = the shortest list, B and C the others
        const int qa = len[0] <= len[1] ? (len[0] <= len[2] ? 0 : 2) : (len[1] <= len[2] ? 1 : 2);
        const int qb = qa == 0 ? 1 : 0, qc = qa == 2 ? 1 : 2;
        const int as = st[qa], an = len[qa];
        const int bs = st[qb], be = st[qb] + len[qb], cs = st[qc], ce = st[qc] + len[qc];
        for (int base = 0; base < an; base += FPH_WAVE) {
            const int ia = base + lane;
            bool ok = false;
            int w = 0;
            if (ia < an && adj_rank[as + ia] < rt) {
                w = adj_idx[as + ia];
                const int pb = bs + lower_bound_from(adj_idx, bs, 0, be, w);
                if (pb < be && adj_idx[pb] == w && adj_rank[pb] < rt) {
                    const int pc = cs + lower_bound_from(adj_idx, cs, 0, ce, w);
                    ok = pc < ce && adj_idx[pc] == w && adj_rank[pc] < rt;
                }
            }
            const unsigned long long okm = __ballot(ok);
            if (okm) {
                found = __shfl(w, __ffsll((long long)okm) - 1);
                break;
            }
        }
    }
    if (lane != 0) return;
    bool app = false;
    if (found >= 0 && found < c) {
        // s = the vertex of t off its largest edge
        int s_vertex = c;
        if (found > a) {
            if (edge_rank(adj_ptr, adj_idx, adj_rank, b, c) == rt) s_vertex = a;
            else if (edge_rank(adj_ptr, adj_idx, adj_rank, a, c) == rt) s_vertex = b;
        }
        app = found < s_vertex;
    }
    long long cx = 0;
    if (app) {
        long long v[4] = {a, b, c, found};
        for (int i = 3; i > 0 && v[i] < v[i - 1]; i--) {
            const long long x = v[i]; v[i] = v[i - 1]; v[i - 1] = x;
        }
        cx = v[3] * (v[3] - 1) / 2 * ((v[3] - 2) * (v[3] - 3) / 2) / 6 +
             v[2] * (v[2] - 1) * (v[2] - 2) / 6 + v[1] * (v[1] - 1) / 2 + v[0];
    }
    flag[t] = app ? 1 : 0;
    colex[t] = cx;
}

}  // namespace

extern "C" int fph_apparent_scan_h2(const int32_t* tri_v0, const int32_t* tri_v1,
                                    const int32_t* tri_v2, const int32_t* tri_max_rank,
                                    int64_t T, const int32_t* adj_ptr, const int32_t* adj_idx,
                                    const int32_t* adj_rank, int64_t max_degree,
                                    int32_t* is_apparent, int64_t* pivot_colex, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (T < 0 || (T > 0 && (!tri_v0 || !tri_v1 || !tri_v2 || !tri_max_rank || !adj_ptr ||
                            !adj_idx || !adj_rank || !is_apparent || !pivot_colex)))
        return FPH_ERR_ARG;
    if (T == 0) return FPH_OK;
    if (T / kScanWavesH2 >= 0x7fffffffLL) return FPH_ERR_LIMIT;
    long long cap = 0;  // max_degree <= 0: no step cap
    if (max_degree > 0) {
        long long md = 64;
        while (md < max_degree && md < 2048) md <<= 1;
        cap = 3 * md;
    }
    k_apparent_scan_h2<<<fph_blocks(T, kScanWavesH2), kScanWavesH2 * FPH_WAVE, 0, s>>>(
        tri_v0, tri_v1, tri_v2, tri_max_rank, T, adj_ptr, adj_idx, adj_rank, cap, is_apparent,
        (long long*)pivot_colex);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

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
