// Host (CPU) Vietoris-Rips pipeline of flash_ph, for point clouds that live
// in host memory -- the reference computes CPU tensors on the CPU
// (rips.py:221-233, persistence.py:197-206).  Same filtration and pairing
// as the gfx950 library, so the diagrams are bit-identical to it:
//   * distances in the direct FP32 form, acc = fma(x_i[k] - x_j[k], ..., acc)
//     over k = 0..d-1, kept iff < fp32(t^2) (kernels/edge_kernel.py:112-125;
//     the reference's CPU route uses cdist's GEMM form instead);
//   * edge order: key float_bits(dsq) << 32 | colex mod 2^32, ties in the
//     emission order of 32 x 32 tiles (rips.py:30-50, 188-202);
//   * minimum spanning forest by rank (Kruskal; = boruvka.py:17-113);
//   * triangles / tetrahedra grouped by their largest edge, the reference's
//     (max_rank, colex) order (reduce_h1_gpu.py:509-521,
//     kernels/apparent_pair_h2_kernel.py:52-58);
//   * apparent pairs with forest clearing (reduce_h1_gpu.py:159-223), the
//     residual Z/2 reduction with the reference's store capacity and MAX_COL
//     (cohomology_general.py:202-371, reduce_h1_gpu.py:282-289,
//     _numba_utils.py:10-104) or the complete reduction, the H1 -> H2
//     clearing (cohomology_general.py:171-173) and the row order of
//     reduce_h1_gpu.py:495-502, 574-604.
// Edge screening, CSR rows, simplex enumeration and cofacet columns run in
// parallel (OpenMP); the reduction is serial, as the reference's.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <numeric>
#include <vector>

#include <omp.h>
#include <parallel/algorithm>

#include "../../include/flashph_cpu.h"

namespace {

using u64 = unsigned long long;

struct Result {
    std::vector<float> dg[3];  // (K, 2) rows per dimension
    int64_t stats[FPHC_NSTATS] = {0};
};

inline float dist_sq(const float* a, const float* b, int d) {
    float acc = 0.0f;
    for (int k = 0; k < d; k++) {
        const float diff = a[k] - b[k];
        acc = std::fmaf(diff, diff, acc);
    }
    return acc;
}

inline uint32_t fbits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

inline float sqrt_rn(float x) { return (float)std::sqrt((double)x); }

// ---------------------------------------------------------------- edges

struct Edge {
    u64 key;
    int32_t i, j;
};

struct Filtration {
    int64_t n = 0, E = 0;
    std::vector<int32_t> ei, ej;
    std::vector<float> dsq, filt;
    std::vector<int64_t> ptr;          // n + 1
    std::vector<int32_t> idx, rank;    // 2E, neighbours ascending per row
};

// Tile rows of 32 points screened in parallel; each tile row's edges are
// kept in the reference's emission order (tile (I, J) row-major, then
// (i mod 32, j mod 32)), and the rows are concatenated in I order.
void screen(const float* pts, int64_t n, int d, float thr, std::vector<Edge>& out) {
    const int64_t nb = (n + 31) / 32;
    std::vector<std::vector<Edge>> rows((size_t)nb);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t I = 0; I < nb; I++) {
        std::vector<Edge>& v = rows[(size_t)I];
        for (int64_t J = I; J < nb; J++) {
            for (int a = 0; a < 32; a++) {
                const int64_t i = I * 32 + a;
                if (i >= n) break;
                const float* xi = pts + i * d;
                for (int b = 0; b < 32; b++) {
                    const int64_t j = J * 32 + b;
                    if (j >= n) break;
                    if (I == J && b <= a) continue;
                    const float q = dist_sq(xi, pts + j * d, d);
                    if (!(q < thr)) continue;
                    const u64 colex = (u64)j * (u64)(j - 1) / 2 + (u64)i;
                    v.push_back({((u64)fbits(q) << 32) | (colex & 0xFFFFFFFFull), (int32_t)i,
                                 (int32_t)j});
                }
            }
        }
    }
    size_t total = 0;
    for (auto& v : rows) total += v.size();
    out.clear();
    out.reserve(total);
    for (auto& v : rows) out.insert(out.end(), v.begin(), v.end());
}

// Sort by key; equal keys (possible only once the colex field wraps,
// n > 92682) keep the emission order, or with exact_colex the order is
// (dsq, j, i) with the full colex index.
void sort_edges(std::vector<Edge>& e, int64_t n, bool exact_colex) {
    if (n <= 92682) {
        __gnu_parallel::sort(e.begin(), e.end(),
                             [](const Edge& a, const Edge& b) { return a.key < b.key; });
    } else if (exact_colex) {
        __gnu_parallel::sort(e.begin(), e.end(), [](const Edge& a, const Edge& b) {
            const uint32_t da = (uint32_t)(a.key >> 32), db = (uint32_t)(b.key >> 32);
            if (da != db) return da < db;
            return a.j != b.j ? a.j < b.j : a.i < b.i;
        });
    } else {
        __gnu_parallel::stable_sort(e.begin(), e.end(),
                                    [](const Edge& a, const Edge& b) { return a.key < b.key; });
    }
}

void build_filtration(const float* pts, int64_t n, int d, float thr, bool exact_colex,
                      Filtration& f) {
    std::vector<Edge> e;
    screen(pts, n, d, thr, e);
    sort_edges(e, n, exact_colex);
    const int64_t E = (int64_t)e.size();
    f.n = n;
    f.E = E;
    f.ei.resize(E);
    f.ej.resize(E);
    f.dsq.resize(E);
    f.filt.resize(E);
#pragma omp parallel for
    for (int64_t k = 0; k < E; k++) {
        f.ei[k] = e[k].i;
        f.ej[k] = e[k].j;
        const uint32_t b = (uint32_t)(e[k].key >> 32);
        float q;
        std::memcpy(&q, &b, 4);
        f.dsq[k] = q;
        f.filt[k] = sqrt_rn(q);
    }
    // CSR (rips.py:53-95): rows by vertex, neighbours ascending
    f.ptr.assign(n + 1, 0);
    for (int64_t k = 0; k < E; k++) {
        f.ptr[f.ei[k] + 1]++;
        f.ptr[f.ej[k] + 1]++;
    }
    for (int64_t v = 0; v < n; v++) f.ptr[v + 1] += f.ptr[v];
    f.idx.resize(2 * E);
    f.rank.resize(2 * E);
    std::vector<int64_t> fill(f.ptr.begin(), f.ptr.end() - 1);
    for (int64_t k = 0; k < E; k++) {
        const int64_t a = fill[f.ei[k]]++, b = fill[f.ej[k]]++;
        f.idx[a] = f.ej[k];
        f.rank[a] = (int32_t)k;
        f.idx[b] = f.ei[k];
        f.rank[b] = (int32_t)k;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t v = 0; v < n; v++) {
        const int64_t lo = f.ptr[v], hi = f.ptr[v + 1];
        std::vector<std::pair<int32_t, int32_t>> row;
        row.reserve(hi - lo);
        for (int64_t h = lo; h < hi; h++) row.emplace_back(f.idx[h], f.rank[h]);
        std::sort(row.begin(), row.end());
        for (int64_t h = lo; h < hi; h++) {
            f.idx[h] = row[h - lo].first;
            f.rank[h] = row[h - lo].second;
        }
    }
}

// rank of edge (s, t), -1 if absent
inline int32_t edge_rank(const Filtration& f, int32_t s, int32_t t) {
    const int32_t* b = f.idx.data() + f.ptr[s];
    const int32_t* e = f.idx.data() + f.ptr[s + 1];
    const int32_t* p = std::lower_bound(b, e, t);
    return (p != e && *p == t) ? f.rank[p - f.idx.data()] : -1;
}

// ------------------------------------------------------- forest and H0

int32_t uf_find(std::vector<int32_t>& p, int32_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

void forest_h0(const Filtration& f, std::vector<uint8_t>& mst, std::vector<float>& h0,
               int64_t* ncomp) {
    mst.assign(std::max<int64_t>(f.E, 1), 0);
    std::vector<int32_t> parent(f.n);
    std::iota(parent.begin(), parent.end(), 0);
    int64_t joined = 0;
    h0.clear();
    for (int64_t k = 0; k < f.E && joined < f.n - 1; k++) {
        const int32_t a = uf_find(parent, f.ei[k]), b = uf_find(parent, f.ej[k]);
        if (a == b) continue;
        parent[std::max(a, b)] = std::min(a, b);
        mst[k] = 1;
        joined++;
        h0.push_back(0.0f);
        h0.push_back(f.filt[k]);  // rank order = ascending length
    }
    *ncomp = f.n - joined;
    for (int64_t c = 0; c < f.n - joined; c++) {
        h0.push_back(0.0f);
        h0.push_back(std::numeric_limits<float>::infinity());
    }
}

// ----------------------------------------------------- simplices / columns

// Triangles grouped by their largest edge e = (u, v), third vertex w
// ascending: global rank E + off[e] + position.
struct Triangles {
    std::vector<int64_t> off;           // E + 1
    std::vector<int32_t> w, e, ru, rv;  // T
    int64_t T = 0;
};

void triangles(const Filtration& f, Triangles& t) {
    const int64_t E = f.E;
    std::vector<int64_t> cnt(E + 1, 0);
    auto each = [&](int64_t e, auto&& emit) {
        const int32_t u = f.ei[e], v = f.ej[e];
        int64_t a = f.ptr[u], b = f.ptr[v];
        const int64_t ae = f.ptr[u + 1], be = f.ptr[v + 1];
        while (a < ae && b < be) {
            const int32_t x = f.idx[a], y = f.idx[b];
            if (x < y) a++;
            else if (y < x) b++;
            else {
                if (f.rank[a] < e && f.rank[b] < e) emit(x, f.rank[a], f.rank[b]);
                a++;
                b++;
            }
        }
    };
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t e = 0; e < E; e++) {
        int64_t c = 0;
        each(e, [&](int32_t, int32_t, int32_t) { c++; });
        cnt[e + 1] = c;
    }
    for (int64_t e = 0; e < E; e++) cnt[e + 1] += cnt[e];
    t.T = cnt[E];
    t.off = cnt;
    t.w.resize(t.T);
    t.e.resize(t.T);
    t.ru.resize(t.T);
    t.rv.resize(t.T);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t e = 0; e < E; e++) {
        int64_t q = t.off[e];
        each(e, [&](int32_t w, int32_t ru, int32_t rv) {
            t.w[q] = w;
            t.e[q] = (int32_t)e;
            t.ru[q] = ru;
            t.rv[q] = rv;
            q++;
        });
    }
}

// local rank of the triangle whose largest edge is m and third vertex x
inline int64_t tri_local(const Triangles& t, int64_t m, int32_t x) {
    const int32_t* b = t.w.data() + t.off[m];
    const int32_t* e = t.w.data() + t.off[m + 1];
    return t.off[m] + (std::lower_bound(b, e, x) - b);
}

// local rank of triangle {s, x, y} from the ranks of (s,x), (s,y), (x,y)
inline int64_t tri_face(const Triangles& t, int32_t s, int32_t x, int32_t y, int32_t rsx,
                        int32_t rsy, int32_t rxy) {
    if (rsx > rsy && rsx > rxy) return tri_local(t, rsx, y);
    if (rsy > rxy) return tri_local(t, rsy, x);
    return tri_local(t, rxy, s);
}

// Columns in CSR form: rows are global ranks, ascending per column.
struct Columns {
    std::vector<int64_t> off;
    std::vector<int32_t> data;
};

// H1 columns: every triangle containing each non-forest edge.
void h1_columns(const Filtration& f, const Triangles& t, const std::vector<uint8_t>& mst,
                Columns& c) {
    const int64_t E = f.E;
    std::vector<std::vector<int32_t>> col((size_t)E);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t e = 0; e < E; e++) {
        if (mst[e]) continue;
        const int32_t u = f.ei[e], v = f.ej[e];
        int64_t a = f.ptr[u], b = f.ptr[v];
        const int64_t ae = f.ptr[u + 1], be = f.ptr[v + 1];
        std::vector<int32_t>& out = col[(size_t)e];
        while (a < ae && b < be) {
            const int32_t x = f.idx[a], y = f.idx[b];
            if (x < y) a++;
            else if (y < x) b++;
            else {
                const int32_t ru = f.rank[a], rv = f.rank[b];
                int64_t loc;
                if (ru < e && rv < e) loc = tri_local(t, e, x);
                else if (ru > rv) loc = tri_local(t, ru, v);
                else loc = tri_local(t, rv, u);
                out.push_back((int32_t)(E + loc));
                a++;
                b++;
            }
        }
        std::sort(out.begin(), out.end());
    }
    c.off.assign(E + 1, 0);
    for (int64_t e = 0; e < E; e++) c.off[e + 1] = c.off[e] + (int64_t)col[(size_t)e].size();
    c.data.resize(c.off[E]);
#pragma omp parallel for
    for (int64_t e = 0; e < E; e++)
        std::copy(col[(size_t)e].begin(), col[(size_t)e].end(), c.data.begin() + c.off[e]);
}

// Tetrahedra grouped by their largest edge e = (u, v): the other two
// vertices x < y are third vertices of e's triangles with (x, y) below e,
// in (y, x) order; key y << 32 | x; global rank E + T + off[e] + position.
struct Tetrahedra {
    std::vector<int64_t> off;
    std::vector<u64> key;
    std::vector<int32_t> e;
    int64_t Q = 0;
};

void tetrahedra(const Filtration& f, const Triangles& t, Tetrahedra& q) {
    const int64_t E = f.E;
    std::vector<int64_t> cnt(E + 1, 0);
    auto each = [&](int64_t e, auto&& emit) {
        const int64_t lo = t.off[e], len = t.off[e + 1] - lo;
        for (int64_t b = 1; b < len; b++)
            for (int64_t a = 0; a < b; a++) {
                const int32_t x = t.w[lo + a], y = t.w[lo + b];
                const int32_t r = edge_rank(f, y, x);
                if (r >= 0 && r < e) emit(x, y);
            }
    };
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t e = 0; e < E; e++) {
        int64_t c = 0;
        each(e, [&](int32_t, int32_t) { c++; });
        cnt[e + 1] = c;
    }
    for (int64_t e = 0; e < E; e++) cnt[e + 1] += cnt[e];
    q.Q = cnt[E];
    q.off = cnt;
    q.key.resize(q.Q);
    q.e.resize(q.Q);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t e = 0; e < E; e++) {
        int64_t p = q.off[e];
        each(e, [&](int32_t x, int32_t y) {
            q.key[p] = ((u64)(uint32_t)y << 32) | (uint32_t)x;
            q.e[p] = (int32_t)e;
            p++;
        });
    }
}

inline int64_t tet_local(const Tetrahedra& q, int64_t m, int32_t p, int32_t r) {
    const u64 k = p > r ? (((u64)(uint32_t)p << 32) | (uint32_t)r)
                        : (((u64)(uint32_t)r << 32) | (uint32_t)p);
    const u64* b = q.key.data() + q.off[m];
    const u64* e = q.key.data() + q.off[m + 1];
    return q.off[m] + (std::lower_bound(b, e, k) - b);
}

// H2 columns: every tetrahedron containing each uncleared triangle.
void h2_columns(const Filtration& f, const Triangles& t, const Tetrahedra& q,
                const std::vector<uint8_t>& cleared, Columns& c) {
    const int64_t E = f.E, T = t.T;
    std::vector<std::vector<int32_t>> col((size_t)T);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t k = 0; k < T; k++) {
        if (cleared[k]) continue;
        const int32_t e = t.e[k], u = f.ei[e], v = f.ej[e], w = t.w[k];
        // common neighbours d of u, v, w: walk u's row, probe v's and w's
        std::vector<int32_t>& out = col[(size_t)k];
        for (int64_t h = f.ptr[u]; h < f.ptr[u + 1]; h++) {
            const int32_t d = f.idx[h];
            const int32_t rvd = edge_rank(f, v, d);
            if (rvd < 0) continue;
            const int32_t rwd = edge_rank(f, w, d);
            if (rwd < 0) continue;
            const int32_t rud = f.rank[h];
            int64_t loc;
            if (rud < e && rvd < e && rwd < e) loc = tet_local(q, e, w, d);
            else if (rud > rvd && rud > rwd) loc = tet_local(q, rud, v, w);
            else if (rvd > rwd) loc = tet_local(q, rvd, u, w);
            else loc = tet_local(q, rwd, u, v);
            out.push_back((int32_t)(E + T + loc));
        }
        std::sort(out.begin(), out.end());
    }
    c.off.assign(T + 1, 0);
    for (int64_t k = 0; k < T; k++) c.off[k + 1] = c.off[k] + (int64_t)col[(size_t)k].size();
    c.data.resize(c.off[T]);
#pragma omp parallel for
    for (int64_t k = 0; k < T; k++)
        std::copy(col[(size_t)k].begin(), col[(size_t)k].end(), c.data.begin() + c.off[k]);
}

// ------------------------------------------------------------ reduction

// Columns 0..K-1 with global ranks base + k; rows are global ranks; row_cmf
// [r - row_base] = the largest non-skipped face of row r (-1: none); r2f over
// all global ranks.  exact: the complete reduction (every reduced column
// stored, no truncation); else the reference's store capacity and MAX_COL.
// ref_cof: the cofacet entries the reference sizes its store from -- every
// column's, skipped ones included (3 per triangle for H1,
// reduce_h1_gpu.py:286-289, 523-534).
struct Reduced {
    std::vector<int32_t> app_b, app_d;   // apparent pairs (persistence > 0)
    std::vector<int32_t> res_b, res_d;   // residual pairs, processing order
    std::vector<int32_t> ess;            // essential columns (global ranks)
    std::vector<int32_t> pivots;         // every pivot row (zero persistence too)
    int64_t xors = 0, drops = 0, truncations = 0;
};

int64_t xor_merge(const int32_t* a, int64_t na, const int32_t* b, int64_t nb, int32_t* out,
                  int64_t lim, int64_t* trunc) {
    int64_t ia = 0, ib = 0, no = 0;
    while (ia < na && ib < nb) {
        if (no >= lim) break;
        if (a[ia] < b[ib]) out[no++] = a[ia++];
        else if (a[ia] > b[ib]) out[no++] = b[ib++];
        else {
            ia++;
            ib++;
        }
    }
    while (ia < na && no < lim) out[no++] = a[ia++];
    while (ib < nb && no < lim) out[no++] = b[ib++];
    if (ia < na || ib < nb) (*trunc)++;
    return no;
}

void reduce(int64_t K, int64_t base, const Columns& c, const std::vector<uint8_t>& skip,
            const std::vector<int32_t>& row_cmf, int64_t row_base, int64_t nrows,
            const std::vector<float>& r2f, bool exact, int64_t max_col, int64_t max_pool,
            int64_t ref_cof, Reduced& out) {
    const int64_t total_cof = ref_cof;
    int64_t cap = std::max<int64_t>(std::max<int64_t>(2 * total_cof, 16 * K), 100000);
    cap = std::min(cap, max_pool);
    const int64_t lim = exact ? std::numeric_limits<int64_t>::max() : max_col;
    std::vector<int64_t> own_at(nrows, -1), own_len(nrows, 0);  // stored column per pivot row
    std::vector<int32_t> store;
    int64_t stored = 0;
    auto fits = [&](int64_t len) {
        return exact || ((int64_t)store.size() + len <= cap && stored < K);
    };
    auto put = [&](int32_t piv, const int32_t* rows, int64_t len) {
        own_at[piv - row_base] = (int64_t)store.size();
        own_len[piv - row_base] = len;
        store.insert(store.end(), rows, rows + len);
        stored++;
    };
    std::vector<uint8_t> apparent(K, 0);
    for (int64_t k = 0; k < K; k++) {
        if (skip[k] || c.off[k + 1] == c.off[k]) continue;
        const int32_t young = c.data[c.off[k]];
        if (row_cmf[young - row_base] == base + k) apparent[k] = 1;
    }
    // apparent pairs, then the store prefilled in ascending column order
    for (int64_t k = 0; k < K; k++) {
        if (!apparent[k]) continue;
        const int32_t young = c.data[c.off[k]];
        out.pivots.push_back(young);
        if (r2f[young] > r2f[base + k]) {
            out.app_b.push_back((int32_t)(base + k));
            out.app_d.push_back(young);
        }
    }
    for (int64_t k = 0; k < K; k++) {
        if (!apparent[k]) continue;
        const int64_t len = c.off[k + 1] - c.off[k];
        if (fits(len)) put(c.data[c.off[k]], c.data.data() + c.off[k], len);
        else out.drops++;
    }
    // residual columns in descending rank order
    std::vector<int32_t> cur, tmp;
    for (int64_t k = K - 1; k >= 0; k--) {
        if (skip[k] || apparent[k]) continue;
        const int64_t n0 = c.off[k + 1] - c.off[k];
        if (n0 == 0) {
            out.ess.push_back((int32_t)(base + k));
            continue;
        }
        int64_t len = n0;
        if (len > lim) {
            len = lim;
            out.truncations++;
        }
        cur.assign(c.data.begin() + c.off[k], c.data.begin() + c.off[k] + len);
        while (len > 0) {
            const int64_t at = own_at[cur[0] - row_base];
            if (at < 0) break;
            const int64_t ol = own_len[cur[0] - row_base];
            tmp.resize(len + ol);
            len = xor_merge(cur.data(), len, store.data() + at, ol, tmp.data(), lim,
                            &out.truncations);
            cur.assign(tmp.begin(), tmp.begin() + len);
            out.xors++;
        }
        if (len > 0) {
            const int32_t piv = cur[0];
            out.pivots.push_back(piv);
            if (fits(len)) put(piv, cur.data(), len);
            else out.drops++;
            if (r2f[piv] > r2f[base + k]) {
                out.res_b.push_back((int32_t)(base + k));
                out.res_d.push_back(piv);
            }
        } else {
            out.ess.push_back((int32_t)(base + k));
        }
    }
}

// rows [apparent ++ residual ++ essential], then stably by birth value
void assemble(const Reduced& r, const std::vector<float>& r2f, std::vector<float>& dg) {
    struct Row { float b, d; };
    std::vector<Row> rows;
    for (size_t k = 0; k < r.app_b.size(); k++) rows.push_back({r2f[r.app_b[k]], r2f[r.app_d[k]]});
    for (size_t k = 0; k < r.res_b.size(); k++) rows.push_back({r2f[r.res_b[k]], r2f[r.res_d[k]]});
    for (int32_t e : r.ess) rows.push_back({r2f[e], std::numeric_limits<float>::infinity()});
    std::stable_sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return a.b < b.b; });
    dg.resize(2 * rows.size());
    for (size_t k = 0; k < rows.size(); k++) {
        dg[2 * k] = rows[k].b;
        dg[2 * k + 1] = rows[k].d;
    }
}

int run(const float* pts, int64_t n, int d, float thr, int max_dim, int exact_h1,
        int exact_colex, int64_t max_col, int64_t max_pool, Result& res) {
    if (n < 2) {
        if (n == 1) res.dg[0] = {0.0f, std::numeric_limits<float>::infinity()};
        return FPHC_OK;
    }
    Filtration f;
    build_filtration(pts, n, d, thr, exact_colex != 0, f);
    const int64_t E = f.E;
    res.stats[FPHC_STAT_E] = E;
    if (E >= 0x7FFFFFFFLL) return FPHC_ERR_LIMIT;
    std::vector<uint8_t> mst;
    int64_t ncomp = 0;
    forest_h0(f, mst, res.dg[0], &ncomp);
    if (max_dim == 0) return FPHC_OK;
    Triangles t;
    triangles(f, t);
    const int64_t T = t.T;
    res.stats[FPHC_STAT_T] = T;
    if (E + T >= 0x7FFFFFFFLL) return FPHC_ERR_LIMIT;
    if (T == 0) {  // reduce_h1_gpu.py:495-502: non-forest edges essential, rank order
        Reduced r;
        for (int64_t e = 0; e < E; e++)
            if (!mst[e]) r.ess.push_back((int32_t)e);
        assemble(r, f.filt, res.dg[1]);
        if (max_dim >= 2) res.dg[2].clear();
        return FPHC_OK;
    }
    Columns c1;
    h1_columns(f, t, mst, c1);
    // rank -> filtration value: edges, then triangles by their largest edge
    std::vector<float> r2f(E + T);
    std::copy(f.filt.begin(), f.filt.end(), r2f.begin());
    for (int64_t k = 0; k < T; k++) r2f[E + k] = f.filt[t.e[k]];
    // largest non-forest face of each triangle
    std::vector<int32_t> cmf1(T);
#pragma omp parallel for
    for (int64_t k = 0; k < T; k++) {
        const int32_t e = t.e[k];
        int32_t m = -1;
        if (!mst[e]) m = e;
        else {
            if (!mst[t.ru[k]]) m = t.ru[k];
            if (!mst[t.rv[k]]) m = std::max(m, t.rv[k]);
        }
        cmf1[k] = m;
    }
    Reduced r1;
    const bool exact1 = exact_h1 || max_dim >= 2;
    reduce(E, 0, c1, mst, cmf1, E, T, r2f, exact1, max_col, max_pool, 3 * T, r1);
    res.stats[FPHC_STAT_XOR1] = r1.xors;
    res.stats[FPHC_STAT_DROPS] = r1.drops;
    res.stats[FPHC_STAT_TRUNC] = r1.truncations;
    assemble(r1, r2f, res.dg[1]);
    if (max_dim < 2) return FPHC_OK;
    // H2: triangles not among the H1 pivots (cohomology_general.py:171-173)
    std::vector<uint8_t> cleared(T, 0);
    for (int32_t p : r1.pivots) cleared[p - E] = 1;
    Tetrahedra q;
    tetrahedra(f, t, q);
    const int64_t Q = q.Q;
    res.stats[FPHC_STAT_Q] = Q;
    if (E + T + Q >= 0x7FFFFFFFLL) return FPHC_ERR_LIMIT;
    Columns c2;
    h2_columns(f, t, q, cleared, c2);
    std::vector<float> r2f2(E + T + Q);
    std::copy(r2f.begin(), r2f.end(), r2f2.begin());
    for (int64_t k = 0; k < Q; k++) r2f2[E + T + k] = f.filt[q.e[k]];
    // largest uncleared face of each tetrahedron (global rank, -1: none)
    std::vector<int32_t> cmf2(Q);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t p = 0; p < E; p++) {
        const int32_t u = f.ei[p], v = f.ej[p];
        for (int64_t k = q.off[p]; k < q.off[p + 1]; k++) {
            const int32_t y = (int32_t)(q.key[k] >> 32), x = (int32_t)(q.key[k] & 0xFFFFFFFFull);
            const int32_t rux = edge_rank(f, u, x), ruy = edge_rank(f, u, y);
            const int32_t rvx = edge_rank(f, v, x), rvy = edge_rank(f, v, y);
            const int32_t rxy = edge_rank(f, x, y);
            const int64_t faces[4] = {tri_local(t, p, x), tri_local(t, p, y),
                                      tri_face(t, u, x, y, rux, ruy, rxy),
                                      tri_face(t, v, x, y, rvx, rvy, rxy)};
            int64_t m = -1;
            for (int64_t fc : faces)
                if (!cleared[fc]) m = std::max(m, fc);
            cmf2[k] = m < 0 ? -1 : (int32_t)(E + m);
        }
    }
    Reduced r2;
    reduce(T, E, c2, cleared, cmf2, E + T, Q, r2f2, true, max_col, max_pool, c2.off[T], r2);
    res.stats[FPHC_STAT_XOR2] = r2.xors;
    assemble(r2, r2f2, res.dg[2]);
    return FPHC_OK;
}

// ---------------------------------------------------------------- apparent-pair scans
//
// The reference's CPU paths of its unwired scans (Numba,
// kernels/apparent_pair_kernel.py:36-96, apparent_pair_h2_kernel.py:61-153):
// complete merges, no step bound.

int64_t colex3(int64_t a, int64_t b, int64_t c) {  // a < b < c
    return c * (c - 1) * (c - 2) / 6 + b * (b - 1) / 2 + a;
}

int64_t colex4(int64_t a, int64_t b, int64_t c, int64_t d) {  // a < b < c < d
    return d * (d - 1) / 2 * ((d - 2) * (d - 3) / 2) / 6 + colex3(a, b, c);
}

// position of w in the sorted list [s, e), or -1
int64_t find_in(const int32_t* idx, int64_t s, int64_t e, int32_t w) {
    const int32_t* p = std::lower_bound(idx + s, idx + e, w);
    return p != idx + e && *p == w ? p - idx : -1;
}

// Triangle (a, b, c), a < b < c, largest edge rank rt: the smallest common
// neighbour w whose edges to a, b, c all rank below rt gives the youngest
// tetrahedron cofacet when one ranks rt at all (colex4 grows with w); the
// triangle is apparent when w also lies below the vertex off its largest
// edge.  Returns the tetrahedron's colex index, or -1.
int64_t scan_h2_one(int32_t a, int32_t b, int32_t c, int32_t rt, const int32_t* ptr,
                    const int32_t* idx, const int32_t* rank) {
    const int32_t v[3] = {a, b, c};
    int q0 = 0;
    for (int q = 1; q < 3; q++)
        if (ptr[v[q] + 1] - ptr[v[q]] < ptr[v[q0] + 1] - ptr[v[q0]]) q0 = q;
    const int q1 = q0 == 0 ? 1 : 0, q2 = q0 == 2 ? 1 : 2;
    int32_t w = -1;
    for (int64_t h = ptr[v[q0]]; h < ptr[v[q0] + 1]; h++) {
        if (rank[h] >= rt) continue;
        const int32_t x = idx[h];
        const int64_t p1 = find_in(idx, ptr[v[q1]], ptr[v[q1] + 1], x);
        if (p1 < 0 || rank[p1] >= rt) continue;
        const int64_t p2 = find_in(idx, ptr[v[q2]], ptr[v[q2] + 1], x);
        if (p2 < 0 || rank[p2] >= rt) continue;
        w = x;
        break;
    }
    if (w < 0 || w > c) return -1;
    int32_t s = c;  // the vertex off the largest edge
    if (w > a) {
        const int64_t pbc = find_in(idx, ptr[b], ptr[b + 1], c);
        const int64_t pac = find_in(idx, ptr[a], ptr[a + 1], c);
        if (pbc >= 0 && rank[pbc] == rt) s = a;
        else if (pac >= 0 && rank[pac] == rt) s = b;
    }
    if (w >= s) return -1;
    int64_t t[4] = {a, b, c, w};
    for (int i = 3; i > 0 && t[i] < t[i - 1]; i--) std::swap(t[i], t[i - 1]);
    return colex4(t[0], t[1], t[2], t[3]);
}

}  // namespace

extern "C" {

int fphc_apparent_scan(const int32_t* edge_i, const int32_t* edge_j, const int32_t* ranks,
                       int64_t K, const int32_t* adj_ptr, const int32_t* adj_idx,
                       const int32_t* adj_rank, int32_t* is_apparent, int64_t* pivot_colex,
                       int threads) {
    if (K < 0 || (K > 0 && (!edge_i || !edge_j || !ranks || !adj_ptr || !adj_idx || !adj_rank ||
                            !is_apparent || !pivot_colex)))
        return FPHC_ERR_ARG;
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < K; k++) {
        const int32_t re = ranks[k];
        int32_t u = edge_i[re], v = edge_j[re];
        if (u > v) std::swap(u, v);
        int64_t iu = adj_ptr[u], iv = adj_ptr[v];
        const int64_t ue = adj_ptr[u + 1], ve = adj_ptr[v + 1];
        int32_t w = -1;
        while (iu < ue && iv < ve) {
            const int32_t nu = adj_idx[iu], nv = adj_idx[iv];
            if (nu == nv) {
                if (adj_rank[iu] < re && adj_rank[iv] < re) {
                    w = nu;
                    break;
                }
                iu++;
                iv++;
            } else if (nu < nv) {
                iu++;
            } else {
                iv++;
            }
        }
        is_apparent[k] = w >= 0;
        if (w < 0) {
            pivot_colex[k] = 0;
        } else {
            int64_t t[3] = {u, v, w};
            for (int i = 2; i > 0 && t[i] < t[i - 1]; i--) std::swap(t[i], t[i - 1]);
            pivot_colex[k] = colex3(t[0], t[1], t[2]);
        }
    }
    if (threads > 0) omp_set_num_threads(prev);
    return FPHC_OK;
}

int fphc_apparent_scan_h2(const int32_t* tri_v0, const int32_t* tri_v1, const int32_t* tri_v2,
                          const int32_t* tri_max_rank, int64_t T, const int32_t* adj_ptr,
                          const int32_t* adj_idx, const int32_t* adj_rank, int32_t* is_apparent,
                          int64_t* pivot_colex, int threads) {
    if (T < 0 || (T > 0 && (!tri_v0 || !tri_v1 || !tri_v2 || !tri_max_rank || !adj_ptr ||
                            !adj_idx || !adj_rank || !is_apparent || !pivot_colex)))
        return FPHC_ERR_ARG;
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < T; t++) {
        const int64_t cx = scan_h2_one(tri_v0[t], tri_v1[t], tri_v2[t], tri_max_rank[t], adj_ptr,
                                       adj_idx, adj_rank);
        is_apparent[t] = cx >= 0;
        pivot_colex[t] = cx >= 0 ? cx : 0;
    }
    if (threads > 0) omp_set_num_threads(prev);
    return FPHC_OK;
}


const char* fphc_version(void) { return "flash_ph cpu 0.1.0 (OpenMP)"; }

int fphc_rips(const float* pts, int64_t n, int d, float thr_sq, int max_dim, int exact_h1,
              int exact_colex, int64_t max_col, int64_t max_pool, int threads,
              fphc_result** out) {
    if (!out || n < 0 || (n > 0 && (!pts || d < 1)) || max_dim < 0 || max_dim > 2)
        return FPHC_ERR_ARG;
    if (n > 0x7FFFFFFFLL) return FPHC_ERR_LIMIT;
    if (max_col <= 0) max_col = 1048576;
    if (max_pool <= 0) max_pool = 100000000;
    Result* r = new (std::nothrow) Result();
    if (!r) return FPHC_ERR_MEM;
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    int rc;
    try {
        rc = run(pts, n, d, thr_sq, max_dim, exact_h1, exact_colex, max_col, max_pool, *r);
    } catch (const std::bad_alloc&) {
        rc = FPHC_ERR_MEM;
    }
    if (threads > 0) omp_set_num_threads(prev);
    if (rc != FPHC_OK) {
        delete r;
        return rc;
    }
    *out = reinterpret_cast<fphc_result*>(r);
    return FPHC_OK;
}

int64_t fphc_rows(const fphc_result* res, int dim) {
    if (!res || dim < 0 || dim > 2) return -1;
    return (int64_t)reinterpret_cast<const Result*>(res)->dg[dim].size() / 2;
}

int fphc_copy(const fphc_result* res, int dim, float* dst) {
    if (!res || dim < 0 || dim > 2 || !dst) return FPHC_ERR_ARG;
    const auto& v = reinterpret_cast<const Result*>(res)->dg[dim];
    std::copy(v.begin(), v.end(), dst);
    return FPHC_OK;
}

int fphc_stats(const fphc_result* res, int64_t* dst) {
    if (!res || !dst) return FPHC_ERR_ARG;
    std::memcpy(dst, reinterpret_cast<const Result*>(res)->stats, sizeof(int64_t) * FPHC_NSTATS);
    return FPHC_OK;
}

void fphc_free(fphc_result* res) { delete reinterpret_cast<Result*>(res); }

int fphc_filtration(const float* pts, int64_t n, int d, float thr_sq, int exact_colex,
                    int threads, fphc_filtration_t** out) {
    if (!out || n < 0 || (n > 0 && (!pts || d < 1))) return FPHC_ERR_ARG;
    if (n > 0x7FFFFFFFLL) return FPHC_ERR_LIMIT;
    Filtration* f = new (std::nothrow) Filtration();
    if (!f) return FPHC_ERR_MEM;
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    int rc = FPHC_OK;
    try {
        if (n >= 2) build_filtration(pts, n, d, thr_sq, exact_colex != 0, *f);
        else {
            f->n = n;
            f->ptr.assign(n + 1, 0);
        }
    } catch (const std::bad_alloc&) {
        rc = FPHC_ERR_MEM;
    }
    if (threads > 0) omp_set_num_threads(prev);
    if (rc == FPHC_OK && f->E >= 0x7FFFFFFFLL) rc = FPHC_ERR_LIMIT;
    if (rc != FPHC_OK) {
        delete f;
        return rc;
    }
    *out = reinterpret_cast<fphc_filtration_t*>(f);
    return FPHC_OK;
}

int64_t fphc_filtration_edges(const fphc_filtration_t* h) {
    return h ? reinterpret_cast<const Filtration*>(h)->E : -1;
}

int fphc_filtration_copy(const fphc_filtration_t* h, int32_t* edge_i, int32_t* edge_j,
                         float* edge_dsq, int32_t* adj_ptr, int32_t* adj_idx, int32_t* adj_rank,
                         float* adj_dsq) {
    if (!h) return FPHC_ERR_ARG;
    const Filtration& f = *reinterpret_cast<const Filtration*>(h);
    if (edge_i) std::copy(f.ei.begin(), f.ei.end(), edge_i);
    if (edge_j) std::copy(f.ej.begin(), f.ej.end(), edge_j);
    if (edge_dsq) std::copy(f.dsq.begin(), f.dsq.end(), edge_dsq);
    if (adj_ptr)
        for (size_t v = 0; v < f.ptr.size(); v++) adj_ptr[v] = (int32_t)f.ptr[v];
    if (adj_idx) std::copy(f.idx.begin(), f.idx.end(), adj_idx);
    if (adj_rank) std::copy(f.rank.begin(), f.rank.end(), adj_rank);
    if (adj_dsq)
        for (size_t h2 = 0; h2 < f.idx.size(); h2++) adj_dsq[h2] = f.dsq[f.rank[h2]];
    return FPHC_OK;
}

void fphc_filtration_free(fphc_filtration_t* h) { delete reinterpret_cast<Filtration*>(h); }

int fphc_knn_kth(const float* pts, int64_t n, int d, int k, float* kth_dsq, int threads) {
    if (n < 2 || !pts || !kth_dsq || d < 1) return FPHC_ERR_ARG;
    k = std::max(1, std::min<int>(k, (int)(n - 1)));
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
    {
        std::vector<float> row(n - 1);
#pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < n; q++) {
            int64_t m = 0;
            for (int64_t j = 0; j < n; j++)
                if (j != q) row[m++] = dist_sq(pts + q * d, pts + j * d, d);
            std::nth_element(row.begin(), row.begin() + (k - 1), row.end());
            kth_dsq[q] = row[k - 1];
        }
    }
    if (threads > 0) omp_set_num_threads(prev);
    return FPHC_OK;
}

int fphc_far_dsq(const float* pts, int64_t n, int d, float* far_dsq, int threads) {
    if (n < 2 || !pts || !far_dsq || d < 1) return FPHC_ERR_ARG;
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < n; q++) {
        float best = 0.0f;
        for (int64_t j = 0; j < n; j++)
            if (j != q) best = std::max(best, dist_sq(pts + q * d, pts + j * d, d));
        far_dsq[q] = best;
    }
    if (threads > 0) omp_set_num_threads(prev);
    return FPHC_OK;
}

}  // extern "C"
