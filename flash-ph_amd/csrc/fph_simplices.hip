// Stages (c) and (d): triangles, tetrahedra, cofacet columns and apparent
// pairs.
//
// Triangles and tetrahedra are enumerated per *maximal edge* in filtration
// order.  For a triangle whose largest-rank edge is e = (u, v), the colex
// order C(c,3)+C(b,2)+a of its sorted vertices equals the ascending order of
// its third vertex w; for a tetrahedron whose largest edge is e, the colex4
// order equals the (y, x) order of its other two vertices x < y.  Hence the
// reference's stable (max_rank, colex) ranking (reduce_h1_gpu.py:509-521,
// apparent_pair_h2_kernel.py:240-249) is produced directly by a count / scan
// / emit over edges, with no global sort:
//     rank(triangle) = E     + tri_off[e] + (index of w in e's lower link)
//     rank(tetra)    = E + T + tet_off[e] + (index of (y, x) in e's list)
// Vertex lists per edge are stored sorted (tri_w: w; tet_key: y << 32 | x),
// so any simplex's rank is found with one binary search.
#include "fph_common.h"

#include <cstdlib>

#include <rocprim/rocprim.hpp>

namespace {

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------
// Triangles (replaces triangle_kernel.py:28-144 and reduce_h1_gpu.py:65-100,
// 509-521).
// ---------------------------------------------------------------------------

// Links of the edges, one wavefront per edge e = (u, v): the common
// neighbours w are found by probing the shorter neighbour list against the
// longer one (staged in LDS when it fits), one probe per lane, and come out
// in ascending w (the probe list's order) through a wave ballot prefix.  A
// thread per edge walked both lists serially, which left long tails on the
// high-degree edges of dense clouds.  Modes:
//   kLinkCount  own[e] = triangles whose largest edge is e (both other edges
//               older), cof[e] = all triangles containing e (0 for forest
//               edges, whose columns H0 clears);
//   kLinkEmit   e's own triangles in ascending w: w, the owning edge, the
//               ranks of (u, w) and (v, w), and cmf = the largest rank among
//               the triangle's non-forest faces (-1 if none) -- what the
//               reference computes by scatter-max (reduce_h1_gpu.py:186-197);
//   kLinkCof    the cofacet column of a non-forest edge: the global rank of
//               every triangle containing it, unsorted (sorted per column
//               afterwards).
// Wavefronts per workgroup: the autotuned ENUM_BLOCK / 64 (dynamic LDS of
// kLinkStage words per wavefront).
constexpr int kLinkStage = 1024;  // staged neighbour-list words per wavefront
enum { kLinkCount = 0, kLinkEmit = 1, kLinkCof = 2 };

// Wavefronts per workgroup of the link kernels for n items.
inline int link_waves(long long n) {
    return std::max(1, fph_tune_value(FPH_TUNE_ENUM_BLOCK, n, 0) / FPH_WAVE);
}

struct LinkOut {
    int* own_cnt;
    int* cof_cnt;
    const int* tri_off;
    int* tri_w;
    int* tri_e;
    int* tri_ru;
    int* tri_rv;
    int* tri_cmf;
    const int* cof_off;
    unsigned* cof_raw;
};

template <int MODE>
__global__ __launch_bounds__(1024) void k_edge_link(
    const int* __restrict__ ei, const int* __restrict__ ej, long long E,
    const int* __restrict__ aptr, const int* __restrict__ aidx, const int* __restrict__ arank,
    const unsigned char* __restrict__ mst, LinkOut o) {
    extern __shared__ int stage[];
    const int wid = threadIdx.x / FPH_WAVE, lane = fph_lane();
    const long long e = (long long)blockIdx.x * (blockDim.x / FPH_WAVE) + wid;
    if (e >= E) return;  // whole wavefront
    const bool me = mst[e] != 0;
    if (MODE == kLinkCof && me) return;
    const int u = ei[e], v = ej[e];
    const int pu = aptr[u], du = aptr[u + 1] - pu, pv = aptr[v], dv = aptr[v + 1] - pv;
    const bool a_is_u = du <= dv;  // probe the shorter list
    const int pa = a_is_u ? pu : pv, da = a_is_u ? du : dv;
    const int pb = a_is_u ? pv : pu, db = a_is_u ? dv : du;
    const bool staged = db <= kLinkStage;
    int* sb = stage + wid * kLinkStage;
    if (staged) {
        for (int i = lane; i < db; i += FPH_WAVE) sb[i] = aidx[pb + i];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    const unsigned long long lt = fph_lanemask_lt();
    int own = 0, cof = 0;
    int pos = 0;
    if (MODE == kLinkEmit) pos = o.tri_off[e];
    if (MODE == kLinkCof) pos = o.cof_off[e];
    for (int base = 0; base < da; base += FPH_WAVE) {
        const int i = base + lane;
        bool hit = false;
        int w = 0, ra = 0, rb = 0;
        if (i < da) {
            w = aidx[pa + i];
            int lo = 0, hi = db;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const int x = staged ? sb[mid] : aidx[pb + mid];
                if (x < w) lo = mid + 1;
                else hi = mid;
            }
            if (lo < db && (staged ? sb[lo] : aidx[pb + lo]) == w) {
                hit = true;
                ra = arank[pa + i];
                rb = arank[pb + lo];
            }
        }
        const int r1 = a_is_u ? ra : rb;  // rank of (u, w)
        const int r2 = a_is_u ? rb : ra;  // rank of (v, w)
        const bool is_own = hit && r1 < e && r2 < e;
        const unsigned long long m_own = __ballot(is_own);
        if (MODE == kLinkCount) {
            own += __popcll(m_own);
            cof += __popcll(__ballot(hit));
        }
        if (MODE == kLinkEmit) {
            if (is_own) {
                const int at = pos + __popcll(m_own & lt);
                o.tri_w[at] = w;
                o.tri_e[at] = (int)e;
                o.tri_ru[at] = r1;
                o.tri_rv[at] = r2;
                int cmf = -1;
                if (!me) cmf = (int)e;
                else {
                    if (!mst[r1]) cmf = r1;
                    if (!mst[r2] && r2 > cmf) cmf = r2;
                }
                o.tri_cmf[at] = cmf;
            }
            pos += __popcll(m_own);
        }
        if (MODE == kLinkCof) {
            const unsigned long long m_hit = __ballot(hit);
            if (hit) {
                long long rank;
                if (is_own) {
                    rank = E + o.tri_off[e] + own + __popcll(m_own & lt);
                } else {
                    const int mx = r1 > r2 ? r1 : r2;
                    const int third = r1 > r2 ? v : u;
                    const int lo = o.tri_off[mx], len = o.tri_off[mx + 1] - lo;
                    rank = E + lo + fph_lower_bound(o.tri_w + lo, len, third);
                }
                o.cof_raw[pos + __popcll(m_hit & lt)] = (unsigned)rank;
            }
            own += __popcll(m_own);
            pos += __popcll(m_hit);
        }
    }
    if (MODE == kLinkCount && lane == 0) {
        o.own_cnt[e] = own;
        o.cof_cnt[e] = me ? 0 : cof;
    }
}

// ---------------------------------------------------------------------------
// Apparent pairs (replaces reduce_h1_gpu.py:159-223; stage (c)).
// kind: 0 cleared, 1 no cofacet (essential), 2 apparent, 3 residual.
// ---------------------------------------------------------------------------

__global__ void k_fill_i32(int* __restrict__ p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// own[0..n) = 0 and *count = 0 in one launch.
__global__ void k_zero_own_count(int* __restrict__ own, long long n, int* __restrict__ count) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) own[i] = 0;
    if (i == 0) *count = 0;
}

__global__ void k_fill_u8(unsigned char* __restrict__ p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_apparent(long long K, long long col_base, const int* __restrict__ off,
                           const unsigned* __restrict__ data, const unsigned char* __restrict__ skip,
                           const int* __restrict__ row_cmf, long long row_base,
                           unsigned char* __restrict__ kind,
                           int* __restrict__ pivot, int* __restrict__ res_list,
                           int* __restrict__ res_count) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int p = -1;
    unsigned char kd = 0;
    if (c < K) {
        if (skip && skip[c]) {
            kd = 0;
        } else if (off[c + 1] == off[c]) {
            kd = 1;
        } else {
            const unsigned y = data[off[c]];
            if ((long long)row_cmf[y - row_base] == col_base + c) {
                kd = 2;
                p = (int)y;
            } else {
                kd = 3;
            }
        }
        kind[c] = kd;
        pivot[c] = p;
    }
    // residual list: one atomic per wavefront (the order is sorted later)
    const bool res = c < K && kd == 3;
    const unsigned long long m = __ballot(res);
    if (!m) return;
    const int first = __ffsll((long long)m) - 1;
    int base = 0;
    if (fph_lane() == first) base = atomicAdd(res_count, __popcll(m));
    base = __shfl(base, first);
    if (res) res_list[base + __popcll(m & fph_lanemask_lt())] = (int)c;
}

// H1 -> H2 clearing: a triangle that is the pivot of an H1 column (apparent
// or reduced, zero persistence included) is not an H2 column
// (cohomology_general.py:171-173).
__global__ void k_clear_pivots(long long K, const unsigned char* __restrict__ kind,
                               const int* __restrict__ pivot, long long row_base,
                               unsigned char* __restrict__ cleared) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const unsigned char kd = kind[c];
    if ((kd == 2 || kd == 3) && pivot[c] >= 0) cleared[pivot[c] - row_base] = 1;
}

// ---------------------------------------------------------------------------
// Tetrahedra per maximal edge (replaces triangle_kernel.py:152-270 and the
// colex4 ranking of apparent_pair_h2_kernel.py:52-58).
// ---------------------------------------------------------------------------

// Tetrahedra per edge e: pairs a < b of e's own triangles (third vertices
// x = tri_w[lo+a] < y = tri_w[lo+b]) whose edge (x, y) is older than e.  One
// wavefront per edge, one pair per lane, pairs in colex order k = b(b-1)/2+a
// (= the (y, x) order of the output): the pair count of an edge is quadratic
// in its triangles, so a thread per edge left the few heavy edges as a long
// serial tail.
__device__ __forceinline__ void pair_of(long long k, int* a, int* b) {
    long long bb = (long long)((1.0 + sqrt(8.0 * (double)k + 1.0)) * 0.5);
    while (bb * (bb - 1) / 2 > k) bb--;
    while ((bb + 1) * bb / 2 <= k) bb++;
    *b = (int)bb;
    *a = (int)(k - bb * (bb - 1) / 2);
}

__global__ __launch_bounds__(1024) void k_tet_count(
    long long E, const int* __restrict__ aptr, const int* __restrict__ aidx,
    const int* __restrict__ arank, const int* __restrict__ tri_off,
    const int* __restrict__ tri_w, int* __restrict__ cnt) {
    const long long e = (long long)blockIdx.x * (blockDim.x / FPH_WAVE) + threadIdx.x / FPH_WAVE;
    if (e >= E) return;
    const int lane = fph_lane();
    const int lo = tri_off[e], len = tri_off[e + 1] - lo;
    const long long np = (long long)len * (len - 1) / 2;
    int c = 0;
    for (long long base = 0; base < np; base += FPH_WAVE) {
        const long long k = base + lane;
        bool ok = false;
        if (k < np) {
            int a, b;
            pair_of(k, &a, &b);
            const int r = fph_edge_rank(aptr, aidx, arank, tri_w[lo + b], tri_w[lo + a]);
            ok = r >= 0 && r < e;
        }
        c += __popcll(__ballot(ok));
    }
    if (lane == 0) cnt[e] = c;
}

__device__ __forceinline__ int tri_rank_local(const int* tri_off, const int* tri_w, int m,
                                              int third) {
    const int lo = tri_off[m];
    return lo + fph_lower_bound(tri_w + lo, tri_off[m + 1] - lo, third);
}

// Local rank of the triangle {s, x, y} given the ranks of its three edges
// (s,x), (s,y), (x,y): find the largest edge and look up the third vertex.
__device__ __forceinline__ int tri_face_local(const int* tri_off, const int* tri_w, int s, int x,
                                              int y, int rsx, int rsy, int rxy) {
    if (rsx > rsy && rsx > rxy) return tri_rank_local(tri_off, tri_w, rsx, y);
    if (rsy > rxy) return tri_rank_local(tri_off, tri_w, rsy, x);
    return tri_rank_local(tri_off, tri_w, rxy, s);
}

__global__ __launch_bounds__(1024) void k_tet_emit(
    long long E, const int* __restrict__ ei, const int* __restrict__ ej,
    const int* __restrict__ aptr, const int* __restrict__ aidx, const int* __restrict__ arank,
    const int* __restrict__ tri_off, const int* __restrict__ tri_w,
    const int* __restrict__ tri_ru, const int* __restrict__ tri_rv,
    const unsigned char* __restrict__ cleared, const int* __restrict__ tet_off,
    unsigned long long* __restrict__ tet_key, int* __restrict__ tet_e, int* __restrict__ tet_cmf) {
    const long long e = (long long)blockIdx.x * (blockDim.x / FPH_WAVE) + threadIdx.x / FPH_WAVE;
    if (e >= E) return;
    const int lane = fph_lane();
    const int lo = tri_off[e], len = tri_off[e + 1] - lo;
    const long long np = (long long)len * (len - 1) / 2;
    if (np == 0) return;
    const int u = ei[e], v = ej[e];
    int q = tet_off[e];
    const unsigned long long lt = fph_lanemask_lt();
    for (long long base = 0; base < np; base += FPH_WAVE) {
        const long long k = base + lane;
        bool ok = false;
        int a = 0, b = 0, x = 0, y = 0, rxy = -1;
        if (k < np) {
            pair_of(k, &a, &b);
            x = tri_w[lo + a];
            y = tri_w[lo + b];
            rxy = fph_edge_rank(aptr, aidx, arank, y, x);
            ok = rxy >= 0 && rxy < e;
        }
        const unsigned long long m = __ballot(ok);
        if (ok) {
            // faces as local triangle ranks, largest first
            const int f_uvy = lo + b, f_uvx = lo + a;
            const int f_uxy = tri_face_local(tri_off, tri_w, u, x, y, tri_ru[lo + a],
                                             tri_ru[lo + b], rxy);
            const int f_vxy = tri_face_local(tri_off, tri_w, v, x, y, tri_rv[lo + a],
                                             tri_rv[lo + b], rxy);
            long long cmf = -1;
            if (!cleared[f_uvy]) cmf = f_uvy;
            else if (!cleared[f_uvx]) cmf = f_uvx;
            else {
                if (!cleared[f_uxy]) cmf = f_uxy;
                if (!cleared[f_vxy] && f_vxy > cmf) cmf = f_vxy;
            }
            const int at = q + __popcll(m & lt);
            tet_key[at] = ((unsigned long long)(unsigned)y << 32) | (unsigned)x;
            tet_e[at] = (int)e;
            tet_cmf[at] = cmf < 0 ? -1 : (int)(E + cmf);
        }
        q += __popcll(m);
    }
}

// Cofacet column of every uncleared triangle {u, v, w} (u, v the ends of its
// largest edge): tetrahedra {u, v, w, d}, d a common neighbour of all three.
// One wavefront per triangle: the shortest of the three neighbour lists is
// probed against the other two (staged in LDS when they fit), one probe per
// lane.  kCount writes the column lengths, kEmit the tetrahedra's global
// ranks (unsorted; sorted per column afterwards).
__device__ __forceinline__ long long tet_rank_in(const int* tet_off,
                                                 const unsigned long long* tet_key, int m, int p,
                                                 int q) {
    const unsigned long long key = p > q
        ? (((unsigned long long)(unsigned)p << 32) | (unsigned)q)
        : (((unsigned long long)(unsigned)q << 32) | (unsigned)p);
    const int lo = tet_off[m];
    return lo + fph_lower_bound(tet_key + lo, tet_off[m + 1] - lo, key);
}

template <int EMIT>
__global__ __launch_bounds__(1024) void k_tri_link(
    long long T, long long E, const int* __restrict__ ei, const int* __restrict__ ej,
    const int* __restrict__ aptr, const int* __restrict__ aidx, const int* __restrict__ arank,
    const int* __restrict__ tri_w, const int* __restrict__ tri_e,
    const unsigned char* __restrict__ cleared, const int* __restrict__ tet_off,
    const unsigned long long* __restrict__ tet_key, int* __restrict__ cnt,
    const int* __restrict__ cof_off,
    unsigned* __restrict__ cof_raw) {
    extern __shared__ int stage[];
    const int wid = threadIdx.x / FPH_WAVE, lane = fph_lane();
    const long long t = (long long)blockIdx.x * (blockDim.x / FPH_WAVE) + wid;
    if (t >= T) return;  // whole wavefront
    if (cleared[t]) {
        if (!EMIT && lane == 0) cnt[t] = 0;
        return;
    }
    const int e = tri_e[t];
    const int u = ei[e], v = ej[e], w = tri_w[t];
    const int pu = aptr[u], du = aptr[u + 1] - pu;
    const int pv = aptr[v], dv = aptr[v + 1] - pv;
    const int pw = aptr[w], dw = aptr[w + 1] - pw;
    // probe list a = the shortest (0 = u, 1 = v, 2 = w); b, c the others in
    // vertex order (staged halves)
    const int a = (du <= dv && du <= dw) ? 0 : (dv <= dw ? 1 : 2);
    const int pa = a == 0 ? pu : (a == 1 ? pv : pw), da = a == 0 ? du : (a == 1 ? dv : dw);
    const int pb = a == 0 ? pv : pu, db = a == 0 ? dv : du;
    const int pc = a == 2 ? pv : pw, dc = a == 2 ? dv : dw;
    const int half = kLinkStage / 2;
    int* sb = stage + wid * kLinkStage;
    int* sc = sb + half;
    const bool stb = db <= half, stc = dc <= half;
    if (stb) for (int i = lane; i < db; i += FPH_WAVE) sb[i] = aidx[pb + i];
    if (stc) for (int i = lane; i < dc; i += FPH_WAVE) sc[i] = aidx[pc + i];
    if (stb || stc) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    const unsigned long long lt = fph_lanemask_lt();
    int pos = EMIT ? cof_off[t] : 0;
    int count = 0;
    const long long base = E + T;
    for (int b0 = 0; b0 < da; b0 += FPH_WAVE) {
        const int i = b0 + lane;
        bool hit = false;
        int d = 0, jb = 0, jc = 0;
        if (i < da) {
            d = aidx[pa + i];
            int lo = 0, hi = db;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((stb ? sb[mid] : aidx[pb + mid]) < d) lo = mid + 1;
                else hi = mid;
            }
            jb = lo;
            if (jb < db && (stb ? sb[jb] : aidx[pb + jb]) == d) {
                lo = 0;
                hi = dc;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((stc ? sc[mid] : aidx[pc + mid]) < d) lo = mid + 1;
                    else hi = mid;
                }
                jc = lo;
                hit = jc < dc && (stc ? sc[jc] : aidx[pc + jc]) == d;
            }
        }
        const unsigned long long m = __ballot(hit);
        if (EMIT && hit) {
            const int ra = arank[pa + i], rb = arank[pb + jb], rc = arank[pc + jc];
            // back to vertex order: b is the first of the other two, c the second
            const int rud = a == 0 ? ra : rb;
            const int rvd = a == 1 ? ra : (a == 0 ? rb : rc);
            const int rwd = a == 2 ? ra : rc;
            long long local;
            if (rud < e && rvd < e && rwd < e) local = tet_rank_in(tet_off, tet_key, e, w, d);
            else if (rud > rvd && rud > rwd) local = tet_rank_in(tet_off, tet_key, rud, v, w);
            else if (rvd > rwd) local = tet_rank_in(tet_off, tet_key, rvd, u, w);
            else local = tet_rank_in(tet_off, tet_key, rwd, u, v);
            cof_raw[pos + __popcll(m & lt)] = (unsigned)(base + local);
        }
        pos += __popcll(m);
        count += __popcll(m);
    }
    if (!EMIT && lane == 0) cnt[t] = count;
}

// counts[0..K) -> off[0..K], total into *total_dev (int64).
int scan_counts(const int* counts, int* off, long long K, long long* total_dev, void* tmp,
                size_t tmp_bytes, hipStream_t s);

__global__ void k_scan_finish(const int* __restrict__ counts, int* __restrict__ off, long long K,
                              long long* __restrict__ total) {
    // off[K] = off[K-1] + counts[K-1]
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const long long tot = K > 0 ? (long long)off[K - 1] + counts[K - 1] : 0;
        off[K] = (int)tot;
        if (total) total[0] = tot;
    }
}

size_t scan_bytes(long long K) {
    size_t tmp = 0;
    rocprim::exclusive_scan(nullptr, tmp, (const int*)nullptr, (int*)nullptr, 0,
                            (size_t)std::max<long long>(K, 1), rocprim::plus<int>(), (hipStream_t)0);
    return align_up(tmp);
}

int scan_counts(const int* counts, int* off, long long K, long long* total_dev, void* tmp,
                size_t tmp_bytes, hipStream_t s) {
    if (K > 0) {
        size_t t = tmp_bytes;
        FPH_CHECK_HIP(rocprim::exclusive_scan(tmp, t, counts, off, 0, (size_t)K,
                                              rocprim::plus<int>(), s));
    }
    k_scan_finish<<<1, 64, 0, s>>>(counts, off, K, (long long*)total_dev);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

// Sorts each column [off[c], off[c+1]) of src ascending into dst (the
// values of one column are distinct), one wavefront per column, in one
// launch: a value's position is the number of smaller values in its column.
// Up to 256 values the column sits in registers (lane l holds values l,
// l + 64, ...) and each value is broadcast by readlane; a longer column is
// ranked 256 values at a time against the whole column read from memory
// (O(L^2 / 256) per column of L values; such columns are rare).  The
// library's segmented sort takes five launches and a host round trip for
// the same work.
constexpr int kColReg = 4;  // values per lane held in registers

__global__ __launch_bounds__(256) void k_colsort(const int* __restrict__ off, long long K,
                                                 const unsigned* __restrict__ src,
                                                 unsigned* __restrict__ dst) {
    const int lane = fph_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x / FPH_WAVE);
    for (long long c = (long long)blockIdx.x * (blockDim.x / FPH_WAVE) + threadIdx.x / FPH_WAVE;
         c < K; c += nw) {
        const int lo = off[c], L = off[c + 1] - lo;
        if (L <= 0) continue;
        for (int c0 = 0; c0 < L; c0 += kColReg * FPH_WAVE) {
            unsigned v[kColReg];
            int rk[kColReg];
#pragma unroll
            for (int q = 0; q < kColReg; q++) {
                const int i = c0 + q * FPH_WAVE + lane;
                v[q] = i < L ? src[lo + i] : 0xFFFFFFFFu;
                rk[q] = 0;
            }
            if (L <= kColReg * FPH_WAVE) {
                for (int j = 0; j < L; j++) {
                    const int qj = j / FPH_WAVE;
                    const unsigned x = (unsigned)__builtin_amdgcn_readlane(
                        (int)(qj == 0 ? v[0] : qj == 1 ? v[1] : qj == 2 ? v[2] : v[3]),
                        j % FPH_WAVE);
#pragma unroll
                    for (int q = 0; q < kColReg; q++) rk[q] += x < v[q];
                }
            } else {
#pragma unroll 8
                for (int j = 0; j < L; j++) {
                    const unsigned x = src[lo + j];
#pragma unroll
                    for (int q = 0; q < kColReg; q++) rk[q] += x < v[q];
                }
            }
#pragma unroll
            for (int q = 0; q < kColReg; q++) {
                const int i = c0 + q * FPH_WAVE + lane;
                if (i < L) dst[lo + rk[q]] = v[q];
            }
        }
    }
}

// The comparison sort costs sum(L^2) over the columns, at least C^2 / K for C
// values in K columns; past kColSortWork of that the library's segmented
// radix sort (linear in C, five launches and a host round trip) is used.
// The environment variable FLASH_PH_COLSORT_WORK replaces the bound (tests
// force either sort with it).
constexpr double kColSortWork = 1073741824.0;

bool colsort_fits(long long C, long long K) {
    const char* env = std::getenv("FLASH_PH_COLSORT_WORK");
    const double work = env ? std::atof(env) : kColSortWork;
    return K > 0 && (double)C * (double)C <= work * (double)K;
}

size_t segsort_bytes(long long C, long long K) {
    if (colsort_fits(C, K)) return 0;
    size_t tmp = 0;
    (void)rocprim::segmented_radix_sort_keys(nullptr, tmp, (const unsigned*)nullptr,
                                             (unsigned*)nullptr,
                                             (unsigned)std::max<long long>(C, 1),
                                             (unsigned)std::max<long long>(K, 1),
                                             (const int*)nullptr, (const int*)nullptr, 0, 32,
                                             (hipStream_t)0);
    return align_up(tmp);
}

int bit_width(unsigned long long x) {
    int b = 0;
    while (x) { b++; x >>= 1; }
    return b;
}

// Sorts the columns of raw into out: one launch of k_colsort, or the
// segmented radix sort over end_bit bits (tmp: segsort_bytes(C, K) bytes).
int sort_columns(const unsigned* raw, unsigned* out, const int* off, long long K, long long C,
                 int end_bit, void* tmp, size_t tmp_bytes, hipStream_t s) {
    if (colsort_fits(C, K)) {
        const long long blocks = std::min<long long>((K + 3) / 4, 4096);
        k_colsort<<<(unsigned)std::max<long long>(blocks, 1), 256, 0, s>>>(off, K, raw, out);
        FPH_LAUNCH_CHECK();
        return FPH_OK;
    }
    size_t t = tmp_bytes;
    FPH_CHECK_HIP(rocprim::segmented_radix_sort_keys(tmp, t, raw, out, (unsigned)C, (unsigned)K,
                                                     off, off + 1, 0, end_bit, s));
    return FPH_OK;
}


}  // namespace

// ----------------------------------------------------------------------------
// C-ABI
// ----------------------------------------------------------------------------

extern "C" int fph_triangle_count(const int32_t* edge_i, const int32_t* edge_j, int64_t E,
                                  const int32_t* adj_ptr, const int32_t* adj_idx,
                                  const int32_t* adj_rank, const uint8_t* mst_mask,
                                  int32_t* tri_off, int32_t* cof_off, int64_t* totals,
                                  void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || !ws_bytes) return FPH_ERR_ARG;
    const size_t tmp = scan_bytes(E);
    const size_t need = tmp + 2 * align_up(4 * (size_t)std::max<int64_t>(E, 1));
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    int* own = (int*)(w + tmp);
    int* cof = (int*)(w + tmp + align_up(4 * (size_t)std::max<int64_t>(E, 1)));
    if (E > 0) {
        LinkOut o{};
        o.own_cnt = own;
        o.cof_cnt = cof;
        const int nw = link_waves(E);
        k_edge_link<kLinkCount><<<fph_blocks(E, nw), nw * FPH_WAVE, nw * kLinkStage * 4, s>>>(
            edge_i, edge_j, E, adj_ptr, adj_idx, adj_rank, mst_mask, o);
        FPH_LAUNCH_CHECK();
    }
    int rc = scan_counts(own, tri_off, E, (long long*)totals, ws, tmp, s);
    if (rc) return rc;
    return scan_counts(cof, cof_off, E, (long long*)totals + 1, ws, tmp, s);
}

extern "C" int fph_triangle_emit(const int32_t* edge_i, const int32_t* edge_j, int64_t E,
                                 const int32_t* adj_ptr, const int32_t* adj_idx,
                                 const int32_t* adj_rank, const uint8_t* mst_mask,
                                 const int32_t* tri_off, int32_t* tri_w, int32_t* tri_e,
                                 int32_t* tri_ru, int32_t* tri_rv, int32_t* tri_cmf,
                                 void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0) return FPH_ERR_ARG;
    if (E == 0) return FPH_OK;
    LinkOut o{};
    o.tri_off = tri_off;
    o.tri_w = tri_w;
    o.tri_e = tri_e;
    o.tri_ru = tri_ru;
    o.tri_rv = tri_rv;
    o.tri_cmf = tri_cmf;
    const int nw = link_waves(E);
    k_edge_link<kLinkEmit><<<fph_blocks(E, nw), nw * FPH_WAVE, nw * kLinkStage * 4, s>>>(
        edge_i, edge_j, E, adj_ptr, adj_idx, adj_rank, mst_mask, o);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_cofacets_h1(const int32_t* edge_i, const int32_t* edge_j, int64_t E,
                               int64_t T, const int32_t* adj_ptr, const int32_t* adj_idx,
                               const int32_t* adj_rank, const uint8_t* mst_mask,
                               const int32_t* tri_off, const int32_t* tri_w,
                               const int32_t* cof_off, int64_t C, uint32_t* cof_data, void* ws,
                               size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || C < 0 || !ws_bytes) return FPH_ERR_ARG;
    if (E + T >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    // [sort temp | the unsorted columns], then sorted into cof_data
    const size_t tmp = segsort_bytes(C, E);
    const size_t need = tmp + align_up(4 * (size_t)std::max<int64_t>(C, 1));
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    if (C == 0) return FPH_OK;
    unsigned* raw = (unsigned*)((char*)ws + tmp);
    LinkOut o{};
    o.tri_off = tri_off;
    o.tri_w = const_cast<int*>(tri_w);  // read only in this mode
    o.cof_off = cof_off;
    o.cof_raw = raw;
    const int nw = link_waves(E);
    k_edge_link<kLinkCof><<<fph_blocks(E, nw), nw * FPH_WAVE, nw * kLinkStage * 4, s>>>(
        edge_i, edge_j, E, adj_ptr, adj_idx, adj_rank, mst_mask, o);
    FPH_LAUNCH_CHECK();
    const int end_bit = std::max(1, bit_width((unsigned long long)(E + T)));
    return sort_columns(raw, (unsigned*)cof_data, cof_off, E, C, end_bit, ws, tmp, s);
}

extern "C" int fph_apparent(int64_t K, int64_t col_base, const int32_t* col_off,
                            const uint32_t* col_data, const uint8_t* skip_mask,
                            const int32_t* row_cmf, int64_t row_base, int64_t nrows,
                            uint8_t* kind, uint64_t* own, int32_t* pivot, int32_t* res_list,
                            int32_t* res_count, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || nrows < 0) return FPH_ERR_ARG;
    // own (2 nrows words) and the residual count cleared together
    k_zero_own_count<<<fph_blocks(std::max<int64_t>(2 * nrows, 1), 256), 256, 0, s>>>(
        (int*)own, 2 * nrows, res_count);
    FPH_LAUNCH_CHECK();
    if (K == 0) return FPH_OK;
    const int tb = fph_tune_value(FPH_TUNE_APPARENT_BLOCK, K, 0);
    k_apparent<<<fph_blocks(K, tb), tb, 0, s>>>(K, col_base, col_off, col_data, skip_mask,
                                                 row_cmf, row_base, kind, pivot, res_list,
                                                 res_count);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_clear_pivots(int64_t K, const uint8_t* kind, const int32_t* pivot,
                                int64_t row_base, int64_t nrows, uint8_t* cleared,
                                void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || nrows < 0) return FPH_ERR_ARG;
    if (nrows > 0) {
        k_fill_u8<<<fph_blocks(nrows, 256), 256, 0, s>>>(cleared, nrows);
        FPH_LAUNCH_CHECK();
    }
    if (K == 0) return FPH_OK;
    const int tb = fph_tune_value(FPH_TUNE_APPARENT_BLOCK, K, 0);
    k_clear_pivots<<<fph_blocks(K, tb), tb, 0, s>>>(K, kind, pivot, row_base, cleared);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_tet_count(int64_t E, const int32_t* adj_ptr, const int32_t* adj_idx,
                             const int32_t* adj_rank, const int32_t* tri_off,
                             const int32_t* tri_w, int32_t* tet_off, int64_t* total, void* ws,
                             size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || !ws_bytes) return FPH_ERR_ARG;
    const size_t tmp = scan_bytes(E);
    const size_t need = tmp + align_up(4 * (size_t)std::max<int64_t>(E, 1));
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    int* cnt = (int*)((char*)ws + tmp);
    if (E > 0) {
        const int nw = link_waves(E);
        k_tet_count<<<fph_blocks(E, nw), nw * FPH_WAVE, 0, s>>>(
            E, adj_ptr, adj_idx, adj_rank, tri_off, tri_w, cnt);
        FPH_LAUNCH_CHECK();
    }
    return scan_counts(cnt, tet_off, E, (long long*)total, ws, tmp, s);
}

extern "C" int fph_tet_emit(int64_t E, const int32_t* edge_i, const int32_t* edge_j,
                            const int32_t* adj_ptr, const int32_t* adj_idx,
                            const int32_t* adj_rank, const int32_t* tri_off,
                            const int32_t* tri_w, const int32_t* tri_ru, const int32_t* tri_rv,
                            const uint8_t* cleared, const int32_t* tet_off, uint64_t* tet_key,
                            int32_t* tet_e, int32_t* tet_cmf, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0) return FPH_ERR_ARG;
    if (E == 0) return FPH_OK;
    const int nw = link_waves(E);
    k_tet_emit<<<fph_blocks(E, nw), nw * FPH_WAVE, 0, s>>>(
        E, edge_i, edge_j, adj_ptr, adj_idx, adj_rank, tri_off, tri_w, tri_ru, tri_rv, cleared,
        tet_off, (unsigned long long*)tet_key, tet_e, tet_cmf);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_cofacets_h2_count(int64_t T, const int32_t* edge_i, const int32_t* edge_j,
                                     const int32_t* adj_ptr, const int32_t* adj_idx,
                                     const int32_t* tri_w, const int32_t* tri_e,
                                     const uint8_t* cleared, int32_t* cof_off, int64_t* total,
                                     void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (T < 0 || !ws_bytes) return FPH_ERR_ARG;
    const size_t tmp = scan_bytes(T);
    const size_t need = tmp + align_up(4 * (size_t)std::max<int64_t>(T, 1));
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    int* cnt = (int*)((char*)ws + tmp);
    if (T > 0) {
        const int nw = link_waves(T);
        k_tri_link<0><<<fph_blocks(T, nw), nw * FPH_WAVE, nw * kLinkStage * 4, s>>>(
            T, 0, edge_i, edge_j, adj_ptr, adj_idx, nullptr, tri_w, tri_e, cleared, nullptr,
            nullptr, cnt, nullptr, nullptr);
        FPH_LAUNCH_CHECK();
    }
    return scan_counts(cnt, cof_off, T, (long long*)total, ws, tmp, s);
}

extern "C" int fph_cofacets_h2(int64_t T, int64_t E, int64_t Q, const int32_t* edge_i,
                               const int32_t* edge_j, const int32_t* adj_ptr,
                               const int32_t* adj_idx, const int32_t* adj_rank,
                               const int32_t* tri_w, const int32_t* tri_e,
                               const uint8_t* cleared, const int32_t* tet_off,
                               const uint64_t* tet_key, const int32_t* cof_off, int64_t C,
                               uint32_t* cof_data, void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (T < 0 || C < 0 || !ws_bytes) return FPH_ERR_ARG;
    if (E + T + Q >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    // [sort temp | the unsorted columns], then sorted into cof_data
    const size_t tmp = segsort_bytes(C, T);
    const size_t need = tmp + align_up(4 * (size_t)std::max<int64_t>(C, 1));
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    if (C == 0) return FPH_OK;
    unsigned* raw = (unsigned*)((char*)ws + tmp);
    const int nw = link_waves(T);
    k_tri_link<1><<<fph_blocks(T, nw), nw * FPH_WAVE, nw * kLinkStage * 4, s>>>(
        T, E, edge_i, edge_j, adj_ptr, adj_idx, adj_rank, tri_w, tri_e, cleared, tet_off,
        (const unsigned long long*)tet_key, nullptr, cof_off, raw);
    FPH_LAUNCH_CHECK();
    const int end_bit = std::max(1, bit_width((unsigned long long)(E + T + Q)));
    return sort_columns(raw, (unsigned*)cof_data, cof_off, T, C, end_bit, ws, tmp, s);
}
