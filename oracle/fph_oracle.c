/*
 * fph_oracle.c -- CPU restatement of flash-ph's exact Vietoris-Rips pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library under flash-ph_amd/csrc.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never calls it.
 *
 * Every stage restates the reference's algorithm in plain C, single-threaded,
 * in the same order of operations, citing the reference file:line it follows
 * (paths relative to the reference's src/flash_ph/):
 *
 *   stage                         reference
 *   ---------------------------   ------------------------------------------
 *   edge enumeration (FMA chain)  kernels/edge_kernel.py:80-149,152-224
 *   sort key + stable sort        rips.py:30-50,188-202
 *   vertex CSR                    rips.py:53-95
 *   Boruvka MST (with its caps)   boruvka.py:17-113
 *   Kruskal MST (cross-check)     -- (unique MST under distinct ranks)
 *   H0 assembly                   persistence.py:169-192
 *   triangle enumeration          _numba_utils.py:111-183 (== Triton emit order)
 *   face ranks / triangle ranks   reduce_h1_gpu.py:65-100,509-521
 *   cofacet CSR                   reduce_h1_gpu.py:107-152
 *   apparent pre-detect           reduce_h1_gpu.py:159-223
 *   residual reduction (v2)       cohomology_general.py:202-371,
 *                                 _numba_utils.py:14-104, reduce_h1_gpu.py:242-303
 *   H1 assembly / row order       reduce_h1_gpu.py:495-502,574-604
 *   tetrahedra                    _numba_utils.py:186-284
 *   tetrahedron order             kernels/apparent_pair_h2_kernel.py:52-58,118
 *   H1 -> H2 clearing cascade     cohomology_general.py:171-173
 *   k-th NN distance (threshold)  threshold.py:133-176 (distance in the FMA
 *                                 form above, not cdist's GEMM form)
 *   enclosing radius              threshold.py:179-200 (same distances)
 *
 * For max_dim=2 the reference hands H1/H2 to giotto-ph (persistence.py:197-206),
 * which is not available here.  H2 is restated with the same cohomology
 * machinery as H1 (columns = triangles not among the H1 pivots, cofacets =
 * tetrahedra in (max_rank, colex4) order); its diagram is pinned as a multiset
 * by the brute-force homology check in tests/, not by giotto's row order.
 *
 * Numerics: compile with -ffp-contract=off.  The only fused operation is the
 * explicit fmaf() of the distance chain.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_COL 1048576LL          /* _numba_utils.py:10 */
#define ORC_MAX_POOL 100000000LL       /* _numba_utils.py:11 */

typedef struct {
    int64_t len;
    int elem;      /* bytes per element */
    void *p;
} Arr;

enum {
    A_EDGE_I, A_EDGE_J, A_EDGE_DSQ,
    A_ADJ_PTR, A_ADJ_IDX, A_ADJ_RANK, A_ADJ_DSQ,
    A_MST_IDX, A_COMP, A_MST_KRUSKAL, A_H0,
    A_TRI_V0, A_TRI_V1, A_TRI_V2, A_TRI_MR, A_TRI_GRANK,
    A_COF_OFF, A_COF_DATA, A_CMF1, A_APPARENT1,
    A_H1_PB, A_H1_PD, A_H1_EB, A_H1_ALLPIV, A_H1,
    A_TET_V0, A_TET_V1, A_TET_V2, A_TET_V3, A_TET_MR, A_TET_GRANK,
    A_COF2_OFF, A_COF2_DATA, A_APPARENT2,
    A_H2_PB, A_H2_PD, A_H2_EB, A_H2,
    A_STATS,
    A_COUNT
};

/* stats slots */
enum { S_E, S_T, S_Q, S_BORUVKA_ROUNDS, S_RES1, S_RES2, S_POOL_DROPS, S_TRUNC,
       S_XOR1, S_XOR2, S_NCOMP, S_STORE1, S_STORED1, S_NSTATS };

typedef struct {
    Arr a[A_COUNT];
    int64_t stats[S_NSTATS];
    int status;
} OrcResult;

static size_t nz(int64_t x) { return x > 0 ? (size_t)x : 1; }

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) abort();
    return p;
}

static void arr_set(OrcResult *r, int id, void *p, int64_t len, int elem) {
    free(r->a[id].p);
    r->a[id].p = p;
    r->a[id].len = len;
    r->a[id].elem = elem;
}

/* ---------------------------------------------------------------------- */
/* Stable sort helpers (merge sort on index arrays keyed by u64).          */
/* ---------------------------------------------------------------------- */

static void msort_idx_rec(int64_t *idx, int64_t *tmp, int64_t lo, int64_t hi,
                          const uint64_t *key) {
    if (hi - lo < 2) return;
    int64_t mid = lo + (hi - lo) / 2;
    msort_idx_rec(idx, tmp, lo, mid, key);
    msort_idx_rec(idx, tmp, mid, hi, key);
    if (key[idx[mid - 1]] <= key[idx[mid]]) return;
    int64_t i = lo, j = mid, k = lo;
    while (i < mid && j < hi) {
        if (key[idx[j]] < key[idx[i]]) tmp[k++] = idx[j++];
        else tmp[k++] = idx[i++];
    }
    while (i < mid) tmp[k++] = idx[i++];
    while (j < hi) tmp[k++] = idx[j++];
    memcpy(idx + lo, tmp + lo, (size_t)(hi - lo) * sizeof(int64_t));
}

/* Stable argsort of u64 keys (torch.argsort(stable=True) semantics). */
static int64_t *stable_argsort_u64(const uint64_t *key, int64_t n) {
    int64_t *idx = xmalloc((size_t)n * sizeof(int64_t));
    int64_t *tmp = xmalloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    msort_idx_rec(idx, tmp, 0, n, key);
    free(tmp);
    return idx;
}

/* Stable argsort of float keys (values >= 0 or +inf; no NaN). */
static int64_t *stable_argsort_f32(const float *v, int64_t n) {
    uint64_t *key = xmalloc((size_t)n * sizeof(uint64_t));
    for (int64_t i = 0; i < n; i++) {
        uint32_t b;
        memcpy(&b, &v[i], 4);
        /* order-preserving map for IEEE floats */
        key[i] = (b & 0x80000000u) ? (uint64_t)(~b) : (uint64_t)(b | 0x80000000u);
    }
    int64_t *idx = stable_argsort_u64(key, n);
    free(key);
    return idx;
}

static float f32_sqrt(float x) {
    /* correctly rounded: sqrtf is IEEE-exact on x86-64 (sqrtss) */
    return sqrtf(x);
}

/* ---------------------------------------------------------------------- */
/* Edge enumeration: edge_kernel.py:80-149.                                */
/* Emission order: tiles (I,J) row-major over n_blocks^2 with I<=J (B=32), */
/* inside a tile row-major over (i mod 32, j mod 32).  Distance is the     */
/* FP32 chain acc = fma(diff, diff, acc), k = 0..d-1, diff = x_i - x_j.     */
/* ---------------------------------------------------------------------- */

static float dist_sq_fma(const float *pts, int d, int64_t i, int64_t j) {
    const float *xi = pts + i * d, *xj = pts + j * d;
    float acc = 0.0f;
    for (int k = 0; k < d; k++) {
        float diff = xi[k] - xj[k];
        acc = fmaf(diff, diff, acc);
    }
    return acc;
}

static int64_t enumerate_edges(const float *pts, int64_t n, int d, float thr_sq,
                               int32_t **oi, int32_t **oj, float **od) {
    const int B = 32;
    int64_t nb = (n + B - 1) / B;
    int64_t cap = 1024, E = 0;
    int32_t *ei = xmalloc(cap * 4), *ej = xmalloc(cap * 4);
    float *ed = xmalloc(cap * 4);
    for (int64_t I = 0; I < nb; I++) {
        for (int64_t J = I; J < nb; J++) {
            for (int a = 0; a < B; a++) {
                int64_t i = I * B + a;
                if (i >= n) break;
                for (int b = 0; b < B; b++) {
                    int64_t j = J * B + b;
                    if (j >= n) break;
                    if (I == J && !(a < b)) continue;
                    float dsq = dist_sq_fma(pts, d, i, j);
                    if (!(dsq < thr_sq)) continue;
                    if (E == cap) {
                        cap *= 2;
                        ei = realloc(ei, cap * 4);
                        ej = realloc(ej, cap * 4);
                        ed = realloc(ed, cap * 4);
                        if (!ei || !ej || !ed) abort();
                    }
                    ei[E] = (int32_t)i;
                    ej[E] = (int32_t)j;
                    ed[E] = dsq;
                    E++;
                }
            }
        }
    }
    *oi = ei; *oj = ej; *od = ed;
    return E;
}

/* rips.py:30-50 -- (float_bits(dsq) << 32) | (colex & 0xFFFFFFFF) */
static uint64_t edge_sort_key(float dsq, int64_t i, int64_t j) {
    uint32_t bits;
    memcpy(&bits, &dsq, 4);
    int64_t colex = j * (j - 1) / 2 + i;
    return ((uint64_t)bits << 32) | ((uint64_t)colex & 0xFFFFFFFFull);
}

/* ---------------------------------------------------------------------- */
/* Boruvka, restated from boruvka.py:17-113 including its round caps.      */
/* ---------------------------------------------------------------------- */

static int flatten(int64_t *comp, int64_t n) {
    int64_t *nw = xmalloc((size_t)n * 8);
    int it;
    for (it = 0; it < 64; it++) {
        int same = 1;
        for (int64_t v = 0; v < n; v++) {
            nw[v] = comp[comp[v]];
            if (nw[v] != comp[v]) same = 0;
        }
        if (same) break;
        memcpy(comp, nw, (size_t)n * 8);
    }
    free(nw);
    return it;
}

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

static int64_t boruvka_ref(const int32_t *ei, const int32_t *ej, int64_t E, int64_t n,
                           int64_t **mst_out, int64_t **comp_out, int64_t *rounds_out) {
    int64_t *comp = xmalloc((size_t)n * 8);
    for (int64_t v = 0; v < n; v++) comp[v] = v;
    uint8_t *in_mst = calloc((size_t)nz(E), 1);
    int64_t *min_key = xmalloc((size_t)n * 8);
    uint8_t *seen = xmalloc((size_t)n);
    int64_t *sel = xmalloc((size_t)(n + 1) * 8);
    const int64_t SENT = INT64_MAX;
    int64_t nn = n < 2 ? 2 : n;
    int max_iters = (int)ceil(log2((double)nn)) + 2;   /* boruvka.py:49 */
    int64_t rounds = 0;
    for (int it = 0; it < max_iters; it++) {
        flatten(comp, n);
        memset(seen, 0, (size_t)n);
        int64_t K = 0;
        for (int64_t v = 0; v < n; v++)
            if (!seen[comp[v]]) { seen[comp[v]] = 1; K++; }
        if (K == 1) break;
        int64_t ncross = 0;
        for (int64_t v = 0; v < n; v++) min_key[v] = SENT;
        for (int64_t e = 0; e < E; e++) {
            int64_t cu = comp[ei[e]], cv = comp[ej[e]];
            if (cu == cv) continue;
            ncross++;
            if (e < min_key[cu]) min_key[cu] = e;
            if (e < min_key[cv]) min_key[cv] = e;
        }
        if (ncross == 0) break;
        rounds++;
        int64_t ns = 0;
        for (int64_t v = 0; v < n; v++)
            if (min_key[v] < SENT) sel[ns++] = min_key[v];
        qsort(sel, (size_t)ns, 8, cmp_i64);
        int64_t nu = 0;
        for (int64_t s = 0; s < ns; s++)
            if (nu == 0 || sel[nu - 1] != sel[s]) sel[nu++] = sel[s];
        for (int64_t s = 0; s < nu; s++) in_mst[sel[s]] = 1;
        /* hook higher root -> lower root via scatter amin (include_self) */
        int64_t *newc = xmalloc((size_t)n * 8);
        memcpy(newc, comp, (size_t)n * 8);
        for (int64_t s = 0; s < nu; s++) {
            int64_t su = comp[ei[sel[s]]], sv = comp[ej[sel[s]]];
            int64_t hf = su > sv ? su : sv, ht = su < sv ? su : sv;
            if (ht < newc[hf]) newc[hf] = ht;
        }
        memcpy(comp, newc, (size_t)n * 8);
        free(newc);
    }
    flatten(comp, n);
    int64_t nm = 0;
    for (int64_t e = 0; e < E; e++) nm += in_mst[e];
    int64_t *mst = xmalloc((size_t)nz(nm) * 8);
    nm = 0;
    for (int64_t e = 0; e < E; e++) if (in_mst[e]) mst[nm++] = e;
    free(in_mst); free(min_key); free(seen); free(sel);
    *mst_out = mst;
    *comp_out = comp;
    *rounds_out = rounds;
    return nm;
}

static int64_t uf_find(int64_t *p, int64_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

/* Kruskal by rank: the unique MST when ranks are distinct. */
static int64_t kruskal(const int32_t *ei, const int32_t *ej, int64_t E, int64_t n,
                       uint8_t *mask) {
    int64_t *p = xmalloc((size_t)n * 8);
    for (int64_t v = 0; v < n; v++) p[v] = v;
    int64_t nm = 0;
    for (int64_t e = 0; e < E; e++) {
        int64_t a = uf_find(p, ei[e]), b = uf_find(p, ej[e]);
        mask[e] = 0;
        if (a != b) { p[a > b ? a : b] = a < b ? a : b; mask[e] = 1; nm++; }
    }
    free(p);
    return nm;
}

/* ---------------------------------------------------------------------- */
/* Residual cohomology reduction, restated from                            */
/* cohomology_general.py:202-371 with the pool prefill of                  */
/* _numba_utils.py:68-104 and the sizing of reduce_h1_gpu.py:278-292.     */
/* The hash map (pivot -> stored column) is a dense table here; the        */
/* reference map has capacity >= 4K > #pivots, so it never drops entries. */
/* ---------------------------------------------------------------------- */

typedef struct {
    int32_t *pb, *pd, *eb, *allpiv;
    int64_t npairs, ness, nallpiv;
    int64_t residual, pool_drops, truncations, xors;
    int64_t store_entries, store_columns;  /* the store's final pool_ptr, num_piv */
} ReduceOut;

static int64_t xor_merge(const int32_t *a, int64_t na, const int32_t *b, int64_t nb,
                         int32_t *out, int64_t *trunc, int64_t lim) {
    int64_t ia = 0, ib = 0, no = 0;
    while (ia < na && ib < nb) {
        if (no >= lim) { (*trunc)++; break; }
        if (a[ia] < b[ib]) out[no++] = a[ia++];
        else if (a[ia] > b[ib]) out[no++] = b[ib++];
        else { ia++; ib++; }
    }
    while (ia < na && no < lim) out[no++] = a[ia++];
    while (ib < nb && no < lim) out[no++] = b[ib++];
    if (ia < na || ib < nb) (*trunc)++;
    return no;
}

/*
 * K columns with global ranks col_rank[k]; cofacet CSR (off, data) sorted
 * ascending; skip[k] (cleared columns); rank_to_filt over all global ranks;
 * cmf[r] = max rank among non-skipped faces of cofacet r (-1 if none).
 * The apparent test and pair extraction follow reduce_h1_gpu.py:159-223;
 * apparent columns are stored first (ascending column index), then the
 * residual columns are reduced in descending rank order.
 */
static void reduce_columns(int64_t K, const int32_t *col_rank, const int32_t *off,
                           const int32_t *data, const uint8_t *skip,
                           const float *r2f, int64_t total, uint8_t *apparent_out,
                           ReduceOut *ro, int exact, int64_t max_col, int64_t max_pool) {
    memset(ro, 0, sizeof(*ro));
    int32_t *cmf = xmalloc((size_t)nz(total) * 4);
    for (int64_t r = 0; r < total; r++) cmf[r] = -1;
    for (int64_t k = 0; k < K; k++) {
        if (skip[k]) continue;
        for (int64_t c = off[k]; c < off[k + 1]; c++)
            if (col_rank[k] > cmf[data[c]]) cmf[data[c]] = col_rank[k];
    }
    int64_t total_cof = off[K];
    int64_t pool_cap = 2 * total_cof;
    if (K * 16 > pool_cap) pool_cap = K * 16;
    if (100000 > pool_cap) pool_cap = 100000;
    if (pool_cap > max_pool) pool_cap = max_pool;
    /* exact: a store without capacity limit (grown on demand) and no column
     * truncation -- the complete reduction (the pairs giotto-ph returns). */
    int64_t store_cap = pool_cap;
    int64_t pool_ptr = 0, num_piv = 0;
    int64_t *piv_of = xmalloc((size_t)nz(total) * 8);   /* rank -> stored idx */
    for (int64_t r = 0; r < total; r++) piv_of[r] = -1;
    int64_t *ps = xmalloc((size_t)nz(K) * 8);
    int64_t *pl = xmalloc((size_t)nz(K) * 8);
    int32_t *pool = xmalloc((size_t)pool_cap * 4);
#define ORC_STORE_FITS(cn_) (exact || pool_ptr + (cn_) <= pool_cap)
#define ORC_STORE_GROW(cn_)                                                   \
    do {                                                                      \
        while (pool_ptr + (cn_) > store_cap) {                                \
            store_cap *= 2;                                                   \
            pool = realloc(pool, (size_t)store_cap * 4);                      \
            if (!pool) abort();                                               \
        }                                                                     \
    } while (0)

    ro->pb = xmalloc((size_t)nz(K) * 4);
    ro->pd = xmalloc((size_t)nz(K) * 4);
    ro->eb = xmalloc((size_t)nz(K) * 4);
    ro->allpiv = xmalloc((size_t)nz(K) * 4);

    /* GPU pre-detect: apparent columns, ascending column index. */
    uint8_t *ext_skip = xmalloc((size_t)nz(K));
    for (int64_t k = 0; k < K; k++) {
        int app = 0;
        if (!skip[k] && off[k + 1] > off[k] && cmf[data[off[k]]] == col_rank[k]) app = 1;
        apparent_out[k] = (uint8_t)app;
        ext_skip[k] = skip[k] || app;
    }
    for (int64_t k = 0; k < K; k++) {
        if (!apparent_out[k]) continue;
        int32_t yc = data[off[k]];
        if (r2f[yc] > r2f[col_rank[k]]) {
            ro->pb[ro->npairs] = col_rank[k];
            ro->pd[ro->npairs] = yc;
            ro->npairs++;
        }
        ro->allpiv[ro->nallpiv++] = yc;
    }
    /* pool prefill (_numba_utils.py:68-104) */
    for (int64_t k = 0; k < K; k++) {
        if (!apparent_out[k]) continue;
        int64_t cn = off[k + 1] - off[k];
        if (cn == 0) continue;
        if (ORC_STORE_FITS(cn) && num_piv < K) {
            ORC_STORE_GROW(cn);
            piv_of[data[off[k]]] = num_piv;
            ps[num_piv] = pool_ptr;
            pl[num_piv] = cn;
            memcpy(pool + pool_ptr, data + off[k], (size_t)cn * 4);
            pool_ptr += cn;
            num_piv++;
        } else {
            ro->pool_drops++;
        }
    }
    /* residual loop, descending rank order (np.argsort(col)[::-1]) */
    int64_t *order = xmalloc((size_t)nz(K) * 8);
    {
        uint64_t *key = xmalloc((size_t)nz(K) * 8);
        for (int64_t k = 0; k < K; k++) key[k] = (uint64_t)(uint32_t)col_rank[k];
        int64_t *asc = stable_argsort_u64(key, K);
        for (int64_t k = 0; k < K; k++) order[k] = asc[K - 1 - k];
        free(asc); free(key);
    }
    int64_t maxrow = 16;
    for (int64_t k = 0; k < K; k++)
        if (off[k + 1] - off[k] > maxrow) maxrow = off[k + 1] - off[k];
    int64_t lim = exact ? INT64_MAX : max_col;
    int64_t wcap = max_col > ORC_MAX_COL ? max_col : ORC_MAX_COL;
    if (exact && total + 1 > wcap) wcap = total + 1;
    int32_t *cur = xmalloc((size_t)wcap * 4), *tmp = xmalloc((size_t)wcap * 4);
    int32_t *carry = xmalloc((size_t)wcap * 4);
    for (int64_t ki = 0; ki < K; ki++) {
        int64_t idx = order[ki];
        if (ext_skip[idx]) continue;
        int32_t r_e = col_rank[idx];
        int64_t cn = off[idx + 1] - off[idx];
        if (cn == 0) { ro->eb[ro->ness++] = r_e; continue; }
        ro->residual++;
        /* safety-net apparent check is dead here: cmf matches ext_skip */
        if (cn > lim) { cn = lim; ro->truncations++; }
        memcpy(cur, data + off[idx], (size_t)cn * 4);
        while (cn > 0) {
            int64_t pi = piv_of[cur[0]];
            if (pi == -1) break;
            memcpy(carry, pool + ps[pi], (size_t)pl[pi] * 4);
            cn = xor_merge(cur, cn, carry, pl[pi], tmp, &ro->truncations, lim);
            memcpy(cur, tmp, (size_t)cn * 4);
            ro->xors++;
        }
        if (cn > 0) {
            int32_t piv = cur[0];
            ro->allpiv[ro->nallpiv++] = piv;
            if (ORC_STORE_FITS(cn) && num_piv < K) {
                ORC_STORE_GROW(cn);
                piv_of[piv] = num_piv;
                ps[num_piv] = pool_ptr;
                pl[num_piv] = cn;
                memcpy(pool + pool_ptr, cur, (size_t)cn * 4);
                pool_ptr += cn;
                num_piv++;
            } else {
                ro->pool_drops++;
            }
            if (r2f[piv] > r2f[r_e]) {
                ro->pb[ro->npairs] = r_e;
                ro->pd[ro->npairs] = piv;
                ro->npairs++;
            }
        } else {
            ro->eb[ro->ness++] = r_e;
        }
    }
    (void)maxrow;
    ro->store_entries = pool_ptr;
    ro->store_columns = num_piv;
    free(cur); free(tmp); free(carry); free(order); free(ext_skip);
    free(cmf); free(piv_of); free(ps); free(pl); free(pool);
#undef ORC_STORE_FITS
#undef ORC_STORE_GROW
}

/* Diagram rows: pairs (in list order) ++ essentials (in list order), then a
 * stable sort by birth value (reduce_h1_gpu.py:586-600). */
static float *assemble_rows(const int32_t *pb, const int32_t *pd, int64_t np,
                            const int32_t *eb, int64_t ne, const float *r2f,
                            int64_t *rows_out) {
    int64_t R = np + ne;
    float *raw = xmalloc((size_t)nz(R) * 8);
    float *birth = xmalloc((size_t)nz(R) * 4);
    for (int64_t k = 0; k < np; k++) {
        raw[2 * k] = r2f[pb[k]];
        raw[2 * k + 1] = r2f[pd[k]];
    }
    for (int64_t k = 0; k < ne; k++) {
        raw[2 * (np + k)] = r2f[eb[k]];
        raw[2 * (np + k) + 1] = INFINITY;
    }
    for (int64_t k = 0; k < R; k++) birth[k] = raw[2 * k];
    int64_t *perm = stable_argsort_f32(birth, R);
    float *out = xmalloc((size_t)nz(R) * 8);
    for (int64_t k = 0; k < R; k++) {
        out[2 * k] = raw[2 * perm[k]];
        out[2 * k + 1] = raw[2 * perm[k] + 1];
    }
    free(raw); free(birth); free(perm);
    *rows_out = R;
    return out;
}

static int csr_lookup(const int32_t *ptr, const int32_t *idx, const int32_t *rank,
                      int32_t s, int32_t t) {
    int32_t lo = ptr[s], hi = ptr[s + 1];
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (idx[mid] < t) lo = mid + 1;
        else if (idx[mid] > t) hi = mid;
        else return rank[mid];
    }
    return -1;
}

/* ---------------------------------------------------------------------- */
/* Full pipeline.                                                          */
/* ---------------------------------------------------------------------- */

/* semantics 0: the reference -- H1 for max_dim <= 1 by its capacity-limited
 * store (reduce_h1_gpu.py:438-604); for max_dim = 2 the reference hands H1
 * and H2 to giotto-ph (persistence.py:197-206), whose complete reduction is
 * restated by the exact store.  semantics 1: exact everywhere. */
/*
 * orc_run_ex: orc_run with the reference's two reduction limits as
 * arguments (max_col rows per column addition, _numba_utils.py:10; max_pool
 * store entries, _numba_utils.py:11; 0 = the reference's values), and
 * optionally starting from a given edge set instead of the O(n^2 d)
 * enumeration: in_E >= 0 edges (in_i < in_j, in_dsq) in any order.  They
 * are first put into the reference's emission order (tile (i/32, j/32)
 * row-major, then (i mod 32, j mod 32), kernels/edge_kernel.py:133-149,
 * 199-210) -- the order the enumeration below produces -- so that the key
 * sort's ties resolve as the reference's.  pts is still needed for
 * nothing else than n; it may be NULL when in_E >= 0.
 */
OrcResult *orc_run_ex(const float *pts, int64_t n, int d, float thr_sq, int max_dim,
                      int semantics, const int32_t *in_i, const int32_t *in_j,
                      const float *in_dsq, int64_t in_E, int64_t max_col,
                      int64_t max_pool) {
    if (max_col <= 0) max_col = ORC_MAX_COL;
    if (max_pool <= 0) max_pool = ORC_MAX_POOL;
    const int exact1 = semantics == 1 || max_dim >= 2;
    OrcResult *r = calloc(1, sizeof(OrcResult));
    if (!r) abort();
    for (int k = 0; k < A_COUNT; k++) arr_set(r, k, xmalloc(1), 0, 4);
    if (n < 2) {
        /* persistence.py:143-157 handled by the caller; keep empties */
        r->status = 0;
        return r;
    }
    /* --- edges: emission order, then stable sort by packed key --- */
    int32_t *ui, *uj; float *ud;
    int64_t E;
    if (in_E >= 0) {
        E = in_E;
        const uint64_t nb = (uint64_t)((n + 31) / 32);
        uint64_t *ek = xmalloc((size_t)nz(E) * 8);
        for (int64_t e = 0; e < E; e++) {
            const uint64_t i = (uint64_t)in_i[e], j = (uint64_t)in_j[e];
            ek[e] = (((i >> 5) * nb + (j >> 5)) << 10) | ((i & 31) << 5) | (j & 31);
        }
        int64_t *ep = stable_argsort_u64(ek, E);
        ui = xmalloc((size_t)nz(E) * 4); uj = xmalloc((size_t)nz(E) * 4);
        ud = xmalloc((size_t)nz(E) * 4);
        for (int64_t e = 0; e < E; e++) {
            ui[e] = in_i[ep[e]]; uj[e] = in_j[ep[e]]; ud[e] = in_dsq[ep[e]];
        }
        free(ek); free(ep);
    } else {
        E = enumerate_edges(pts, n, d, thr_sq, &ui, &uj, &ud);
    }
    r->stats[S_E] = E;
    uint64_t *key = xmalloc((size_t)nz(E) * 8);
    for (int64_t e = 0; e < E; e++) key[e] = edge_sort_key(ud[e], ui[e], uj[e]);
    int64_t *perm = stable_argsort_u64(key, E);
    int32_t *ei = xmalloc((size_t)nz(E) * 4), *ej = xmalloc((size_t)nz(E) * 4);
    float *ed = xmalloc((size_t)nz(E) * 4);
    for (int64_t e = 0; e < E; e++) { ei[e] = ui[perm[e]]; ej[e] = uj[perm[e]]; ed[e] = ud[perm[e]]; }
    free(ui); free(uj); free(ud); free(key); free(perm);
    arr_set(r, A_EDGE_I, ei, E, 4);
    arr_set(r, A_EDGE_J, ej, E, 4);
    arr_set(r, A_EDGE_DSQ, ed, E, 4);

    /* --- CSR: 2E half edges stably sorted by src*n+dst (rips.py:74-95) --- */
    int32_t *aptr = calloc((size_t)n + 1, 4);
    int32_t *aidx = xmalloc((size_t)nz(2 * E) * 4);
    int32_t *arank = xmalloc((size_t)nz(2 * E) * 4);
    float *adsq = xmalloc((size_t)nz(2 * E) * 4);
    {
        uint64_t *hk = xmalloc((size_t)nz(2 * E) * 8);
        for (int64_t e = 0; e < E; e++) {
            hk[e] = (uint64_t)ei[e] * (uint64_t)n + (uint64_t)ej[e];
            hk[E + e] = (uint64_t)ej[e] * (uint64_t)n + (uint64_t)ei[e];
        }
        int64_t *hp = stable_argsort_u64(hk, 2 * E);
        for (int64_t h = 0; h < 2 * E; h++) {
            int64_t s = hp[h];
            int64_t e = s < E ? s : s - E;
            int32_t src = s < E ? ei[e] : ej[e];
            int32_t dst = s < E ? ej[e] : ei[e];
            aidx[h] = dst;
            arank[h] = (int32_t)e;
            adsq[h] = ed[e];
            aptr[src + 1]++;
        }
        for (int64_t v = 0; v < n; v++) aptr[v + 1] += aptr[v];
        free(hk); free(hp);
    }
    arr_set(r, A_ADJ_PTR, aptr, n + 1, 4);
    arr_set(r, A_ADJ_IDX, aidx, 2 * E, 4);
    arr_set(r, A_ADJ_RANK, arank, 2 * E, 4);
    arr_set(r, A_ADJ_DSQ, adsq, 2 * E, 4);

    /* --- MST --- */
    int64_t *mst, *comp, rounds;
    int64_t nm = boruvka_ref(ei, ej, E, n, &mst, &comp, &rounds);
    r->stats[S_BORUVKA_ROUNDS] = rounds;
    arr_set(r, A_MST_IDX, mst, nm, 8);
    arr_set(r, A_COMP, comp, n, 8);
    uint8_t *kmask = xmalloc((size_t)nz(E));
    kruskal(ei, ej, E, n, kmask);
    arr_set(r, A_MST_KRUSKAL, kmask, E, 1);
    uint8_t *mst_mask = calloc((size_t)nz(E), 1);
    for (int64_t k = 0; k < nm; k++) mst_mask[mst[k]] = 1;

    /* --- H0 (persistence.py:169-192) --- */
    {
        uint8_t *seen = calloc((size_t)n, 1);
        int64_t ncomp = 0;
        for (int64_t v = 0; v < n; v++) if (!seen[comp[v]]) { seen[comp[v]] = 1; ncomp++; }
        free(seen);
        r->stats[S_NCOMP] = ncomp;
        float *h0 = xmalloc((size_t)(nm + ncomp) * 8);
        float *vals = xmalloc((size_t)nz(nm) * 4);
        for (int64_t k = 0; k < nm; k++) vals[k] = f32_sqrt(ed[mst[k]]);
        int64_t *pv = stable_argsort_f32(vals, nm);
        for (int64_t k = 0; k < nm; k++) { h0[2 * k] = 0.0f; h0[2 * k + 1] = vals[pv[k]]; }
        for (int64_t k = 0; k < ncomp; k++) { h0[2 * (nm + k)] = 0.0f; h0[2 * (nm + k) + 1] = INFINITY; }
        free(vals); free(pv);
        arr_set(r, A_H0, h0, nm + ncomp, 8);
    }
    if (max_dim == 0) { free(mst_mask); return r; }

    float *edge_filt = xmalloc((size_t)nz(E) * 4);
    for (int64_t e = 0; e < E; e++) edge_filt[e] = f32_sqrt(ed[e]);

    /* --- triangles: per edge, common neighbours w > v in ascending w --- */
    int64_t T = 0, tcap = 1024;
    int32_t *t0 = xmalloc(tcap * 4), *t1 = xmalloc(tcap * 4), *t2 = xmalloc(tcap * 4),
            *tm = xmalloc(tcap * 4);
    for (int64_t e = 0; e < E; e++) {
        int32_t u = ei[e], v = ej[e];
        int32_t iu = aptr[u], iv = aptr[v];
        while (iu < aptr[u + 1] && iv < aptr[v + 1]) {
            int32_t nu = aidx[iu], nv = aidx[iv];
            if (nu == nv) {
                if (nu > v) {
                    int32_t mr = (int32_t)e;
                    if (arank[iu] > mr) mr = arank[iu];
                    if (arank[iv] > mr) mr = arank[iv];
                    if (T == tcap) {
                        tcap *= 2;
                        t0 = realloc(t0, tcap * 4); t1 = realloc(t1, tcap * 4);
                        t2 = realloc(t2, tcap * 4); tm = realloc(tm, tcap * 4);
                        if (!t0 || !t1 || !t2 || !tm) abort();
                    }
                    t0[T] = u; t1[T] = v; t2[T] = nu; tm[T] = mr; T++;
                }
                iu++; iv++;
            } else if (nu < nv) iu++;
            else iv++;
        }
    }
    r->stats[S_T] = T;
    arr_set(r, A_TRI_V0, t0, T, 4);
    arr_set(r, A_TRI_V1, t1, T, 4);
    arr_set(r, A_TRI_V2, t2, T, 4);
    arr_set(r, A_TRI_MR, tm, T, 4);

    if (T == 0) {
        /* reduce_h1_gpu.py:495-502: all non-MST edges essential, rank order */
        int64_t ne = 0;
        int32_t *eb = xmalloc((size_t)nz(E) * 4);
        for (int64_t e = 0; e < E; e++) if (!mst_mask[e]) eb[ne++] = (int32_t)e;
        int64_t R;
        float *h1 = assemble_rows(NULL, NULL, 0, eb, ne, edge_filt, &R);
        arr_set(r, A_H1, h1, R, 8);
        arr_set(r, A_H1_EB, eb, ne, 4);
        free(edge_filt); free(mst_mask);
        return r;
    }

    /* --- triangle global ranks: stable by colex3, then stable by max rank --- */
    int32_t *tgr = xmalloc((size_t)T * 4);
    int64_t *tri_order = NULL;
    {
        uint64_t *ck = xmalloc((size_t)T * 8);
        for (int64_t t = 0; t < T; t++) {
            int64_t a = t0[t], b = t1[t], c = t2[t];
            ck[t] = (uint64_t)(c * (c - 1) * (c - 2) / 6 + b * (b - 1) / 2 + a);
        }
        int64_t *p1 = stable_argsort_u64(ck, T);
        uint64_t *mk = xmalloc((size_t)T * 8);
        for (int64_t k = 0; k < T; k++) mk[k] = (uint64_t)tm[p1[k]];
        int64_t *p2 = stable_argsort_u64(mk, T);
        tri_order = xmalloc((size_t)T * 8);
        for (int64_t k = 0; k < T; k++) tri_order[k] = p1[p2[k]];
        for (int64_t k = 0; k < T; k++) tgr[tri_order[k]] = (int32_t)(E + k);
        free(ck); free(p1); free(mk); free(p2);
    }
    arr_set(r, A_TRI_GRANK, tgr, T, 4);

    /* --- face ranks (ascending per triangle) and cofacet CSR --- */
    int32_t *coff = calloc((size_t)E + 1, 4);
    int32_t *cdat = xmalloc((size_t)3 * T * 4);
    {
        int32_t *fr = xmalloc((size_t)3 * T * 4);
        for (int64_t t = 0; t < T; t++) {
            fr[3 * t] = csr_lookup(aptr, aidx, arank, t0[t], t1[t]);
            fr[3 * t + 1] = csr_lookup(aptr, aidx, arank, t0[t], t2[t]);
            fr[3 * t + 2] = csr_lookup(aptr, aidx, arank, t1[t], t2[t]);
            coff[fr[3 * t] + 1]++; coff[fr[3 * t + 1] + 1]++; coff[fr[3 * t + 2] + 1]++;
        }
        for (int64_t e = 0; e < E; e++) coff[e + 1] += coff[e];
        /* primary = edge rank, secondary = triangle global rank (both stable) */
        uint64_t *k2 = xmalloc((size_t)3 * T * 8);
        for (int64_t x = 0; x < 3 * T; x++)
            k2[x] = ((uint64_t)(uint32_t)fr[x] << 32) | (uint32_t)tgr[x / 3];
        int64_t *p = stable_argsort_u64(k2, 3 * T);
        for (int64_t x = 0; x < 3 * T; x++) cdat[x] = tgr[p[x] / 3];
        free(k2); free(p); free(fr);
    }
    arr_set(r, A_COF_OFF, coff, E + 1, 4);
    arr_set(r, A_COF_DATA, cdat, 3 * T, 4);

    /* --- rank_to_filt = [edge_filt ; edge_filt[tri max rank, sorted]] --- */
    int64_t total1 = E + T;
    float *r2f = xmalloc((size_t)total1 * 4);
    memcpy(r2f, edge_filt, (size_t)E * 4);
    for (int64_t k = 0; k < T; k++) r2f[E + k] = edge_filt[tm[tri_order[k]]];

    int32_t *colr = xmalloc((size_t)E * 4);
    for (int64_t e = 0; e < E; e++) colr[e] = (int32_t)e;
    uint8_t *app1 = xmalloc((size_t)E);
    ReduceOut ro1;
    reduce_columns(E, colr, coff, cdat, mst_mask, r2f, total1, app1, &ro1, exact1, max_col,
                   max_pool);
    arr_set(r, A_APPARENT1, app1, E, 1);
    r->stats[S_RES1] = ro1.residual;
    r->stats[S_XOR1] = ro1.xors;
    r->stats[S_STORE1] = ro1.store_entries;
    r->stats[S_STORED1] = ro1.store_columns;
    r->stats[S_POOL_DROPS] += ro1.pool_drops;
    r->stats[S_TRUNC] += ro1.truncations;
    {
        int64_t R;
        float *h1 = assemble_rows(ro1.pb, ro1.pd, ro1.npairs, ro1.eb, ro1.ness, r2f, &R);
        arr_set(r, A_H1, h1, R, 8);
    }
    arr_set(r, A_H1_PB, ro1.pb, ro1.npairs, 4);
    arr_set(r, A_H1_PD, ro1.pd, ro1.npairs, 4);
    arr_set(r, A_H1_EB, ro1.eb, ro1.ness, 4);
    arr_set(r, A_H1_ALLPIV, ro1.allpiv, ro1.nallpiv, 4);

    if (max_dim < 2) {
        free(colr); free(r2f); free(edge_filt); free(mst_mask); free(tri_order);
        return r;
    }

    /* --- tetrahedra per triangle (emission order of _numba_utils.py:186-284) --- */
    int64_t Q = 0, qcap = 1024;
    int32_t *q0 = xmalloc(qcap * 4), *q1 = xmalloc(qcap * 4), *q2 = xmalloc(qcap * 4),
            *q3 = xmalloc(qcap * 4), *qm = xmalloc(qcap * 4);
    for (int64_t t = 0; t < T; t++) {
        int32_t a = t0[t], b = t1[t], c = t2[t];
        int32_t ia = aptr[a], ib = aptr[b], ic = aptr[c];
        while (ia < aptr[a + 1] && ib < aptr[b + 1] && ic < aptr[c + 1]) {
            int32_t na = aidx[ia], nbv = aidx[ib], nc = aidx[ic];
            int32_t mx = na;
            if (nbv > mx) mx = nbv;
            if (nc > mx) mx = nc;
            if (na == nbv && nbv == nc) {
                if (na > c) {
                    int32_t mr = tm[t];
                    if (arank[ia] > mr) mr = arank[ia];
                    if (arank[ib] > mr) mr = arank[ib];
                    if (arank[ic] > mr) mr = arank[ic];
                    if (Q == qcap) {
                        qcap *= 2;
                        q0 = realloc(q0, qcap * 4); q1 = realloc(q1, qcap * 4);
                        q2 = realloc(q2, qcap * 4); q3 = realloc(q3, qcap * 4);
                        qm = realloc(qm, qcap * 4);
                        if (!q0 || !q1 || !q2 || !q3 || !qm) abort();
                    }
                    q0[Q] = a; q1[Q] = b; q2[Q] = c; q3[Q] = na; qm[Q] = mr; Q++;
                }
                ia++; ib++; ic++;
            } else {
                if (na < mx) ia++;
                if (nbv < mx) ib++;
                if (nc < mx) ic++;
            }
        }
    }
    r->stats[S_Q] = Q;
    arr_set(r, A_TET_V0, q0, Q, 4);
    arr_set(r, A_TET_V1, q1, Q, 4);
    arr_set(r, A_TET_V2, q2, Q, 4);
    arr_set(r, A_TET_V3, q3, Q, 4);
    arr_set(r, A_TET_MR, qm, Q, 4);

    /* --- tetrahedron order (max_rank, colex4); global rank = E + T + pos --- */
    int32_t *qgr = xmalloc((size_t)nz(Q) * 4);
    int64_t *tet_order = xmalloc((size_t)nz(Q) * 8);
    {
        uint64_t *ck = xmalloc((size_t)nz(Q) * 8);
        for (int64_t q = 0; q < Q; q++) {
            int64_t a = q0[q], b = q1[q], c = q2[q], dd = q3[q];
            int64_t hi = dd * (dd - 1) / 2, lo = (dd - 2) * (dd - 3) / 2;
            ck[q] = (uint64_t)(hi * lo / 6 + c * (c - 1) * (c - 2) / 6 + b * (b - 1) / 2 + a);
        }
        int64_t *p1 = stable_argsort_u64(ck, Q);
        uint64_t *mk = xmalloc((size_t)nz(Q) * 8);
        for (int64_t k = 0; k < Q; k++) mk[k] = (uint64_t)qm[p1[k]];
        int64_t *p2 = stable_argsort_u64(mk, Q);
        for (int64_t k = 0; k < Q; k++) tet_order[k] = p1[p2[k]];
        for (int64_t k = 0; k < Q; k++) qgr[tet_order[k]] = (int32_t)(E + T + k);
        free(ck); free(p1); free(mk); free(p2);
    }
    arr_set(r, A_TET_GRANK, qgr, Q, 4);

    /* triangle global rank from sorted vertex triple (colex lookup) */
    /* triangles sharing a max edge are few; use a hash over colex3 via sort */
    uint64_t *tri_ck = xmalloc((size_t)T * 8);
    int64_t *tri_ck_perm;
    {
        for (int64_t t = 0; t < T; t++) {
            int64_t a = t0[t], b = t1[t], c = t2[t];
            tri_ck[t] = (uint64_t)(c * (c - 1) * (c - 2) / 6 + b * (b - 1) / 2 + a);
        }
        tri_ck_perm = stable_argsort_u64(tri_ck, T);
    }
#define TRI_RANK_OF(A, B, C, OUT) do {                                         \
        int64_t a_ = (A), b_ = (B), c_ = (C);                                   \
        uint64_t k_ = (uint64_t)(c_ * (c_ - 1) * (c_ - 2) / 6 + b_ * (b_ - 1) / 2 + a_); \
        int64_t lo_ = 0, hi_ = T;                                               \
        while (lo_ < hi_) {                                                     \
            int64_t m_ = (lo_ + hi_) / 2;                                       \
            if (tri_ck[tri_ck_perm[m_]] < k_) lo_ = m_ + 1; else hi_ = m_;      \
        }                                                                       \
        (OUT) = tgr[tri_ck_perm[lo_]];                                          \
    } while (0)

    /* cofacet CSR for triangle columns (local index = global rank - E) */
    int32_t *c2off = calloc((size_t)T + 1, 4);
    int32_t *c2dat = xmalloc((size_t)nz(4 * Q) * 4);
    {
        int32_t *fr = xmalloc((size_t)nz(4 * Q) * 4);
        for (int64_t q = 0; q < Q; q++) {
            int32_t a = q0[q], b = q1[q], c = q2[q], dd = q3[q];
            int32_t g;
            TRI_RANK_OF(a, b, c, g); fr[4 * q] = g - (int32_t)E;
            TRI_RANK_OF(a, b, dd, g); fr[4 * q + 1] = g - (int32_t)E;
            TRI_RANK_OF(a, c, dd, g); fr[4 * q + 2] = g - (int32_t)E;
            TRI_RANK_OF(b, c, dd, g); fr[4 * q + 3] = g - (int32_t)E;
            for (int f = 0; f < 4; f++) c2off[fr[4 * q + f] + 1]++;
        }
        for (int64_t t = 0; t < T; t++) c2off[t + 1] += c2off[t];
        uint64_t *k2 = xmalloc((size_t)nz(4 * Q) * 8);
        for (int64_t x = 0; x < 4 * Q; x++)
            k2[x] = ((uint64_t)(uint32_t)fr[x] << 32) | (uint32_t)qgr[x / 4];
        int64_t *p = stable_argsort_u64(k2, 4 * Q);
        for (int64_t x = 0; x < 4 * Q; x++) c2dat[x] = qgr[p[x] / 4];
        free(k2); free(p); free(fr);
    }
#undef TRI_RANK_OF
    arr_set(r, A_COF2_OFF, c2off, T + 1, 4);
    arr_set(r, A_COF2_DATA, c2dat, 4 * Q, 4);

    /* H2 columns: triangles in global-rank order; cleared = H1 pivots */
    int64_t total2 = E + T + Q;
    float *r2f2 = xmalloc((size_t)total2 * 4);
    memcpy(r2f2, r2f, (size_t)(E + T) * 4);
    for (int64_t k = 0; k < Q; k++) r2f2[E + T + k] = edge_filt[qm[tet_order[k]]];
    int32_t *col2 = xmalloc((size_t)T * 4);
    for (int64_t t = 0; t < T; t++) col2[t] = (int32_t)(E + t);
    uint8_t *skip2 = calloc((size_t)T, 1);
    for (int64_t k = 0; k < ro1.nallpiv; k++) skip2[ro1.allpiv[k] - E] = 1;
    /* CSR rows are indexed by local triangle rank; reorder data accordingly */
    uint8_t *app2 = xmalloc((size_t)T);
    ReduceOut ro2;
    reduce_columns(T, col2, c2off, c2dat, skip2, r2f2, total2, app2, &ro2, 1, max_col,
                   max_pool);
    arr_set(r, A_APPARENT2, app2, T, 1);
    r->stats[S_RES2] = ro2.residual;
    r->stats[S_XOR2] = ro2.xors;
    r->stats[S_POOL_DROPS] += ro2.pool_drops;
    r->stats[S_TRUNC] += ro2.truncations;
    {
        int64_t R;
        float *h2 = assemble_rows(ro2.pb, ro2.pd, ro2.npairs, ro2.eb, ro2.ness, r2f2, &R);
        arr_set(r, A_H2, h2, R, 8);
    }
    arr_set(r, A_H2_PB, ro2.pb, ro2.npairs, 4);
    arr_set(r, A_H2_PD, ro2.pd, ro2.npairs, 4);
    arr_set(r, A_H2_EB, ro2.eb, ro2.ness, 4);
    free(ro2.allpiv);
    free(col2); free(skip2); free(r2f2); free(tri_ck); free(tri_ck_perm);
    free(tet_order);
    free(colr); free(r2f); free(edge_filt); free(mst_mask); free(tri_order);
    return r;
}

OrcResult *orc_run(const float *pts, int64_t n, int d, float thr_sq, int max_dim,
                   int semantics) {
    return orc_run_ex(pts, n, d, thr_sq, max_dim, semantics, NULL, NULL, NULL, -1, 0, 0);
}

/* Neighbours of vertex v (every u != v with FMA-chain dsq(min, max) < thr_sq,
 * kernels/edge_kernel.py:112-125), ascending, with their dsq; returns the
 * count (at most cap written).  For membership checks of one vertex's row
 * on clouds too large for the full enumeration. */
int64_t orc_vertex_neighbours(const float *pts, int64_t n, int d, float thr_sq, int64_t v,
                              int32_t *out_idx, float *out_dsq, int64_t cap) {
    int64_t m = 0;
    for (int64_t u = 0; u < n; u++) {
        if (u == v) continue;
        const float dsq = u < v ? dist_sq_fma(pts, d, u, v) : dist_sq_fma(pts, d, v, u);
        if (!(dsq < thr_sq)) continue;
        if (m < cap) { out_idx[m] = (int32_t)u; out_dsq[m] = dsq; }
        m++;
    }
    return m;
}

int64_t orc_len(const OrcResult *r, int id) {
    if (id < 0 || id >= A_COUNT) return -1;
    if (id == A_STATS) return S_NSTATS;
    return r->a[id].len;
}

int orc_elem(const OrcResult *r, int id) {
    if (id == A_STATS) return 8;
    return r->a[id].elem;
}

int orc_get(const OrcResult *r, int id, void *dst) {
    if (id < 0 || id >= A_COUNT) return -1;
    if (id == A_STATS) { memcpy(dst, r->stats, sizeof(r->stats)); return 0; }
    int mult = (id == A_H0 || id == A_H1 || id == A_H2) ? 1 : 1;
    memcpy(dst, r->a[id].p, (size_t)r->a[id].len * r->a[id].elem * mult);
    return 0;
}

void orc_free(OrcResult *r) {
    if (!r) return;
    for (int k = 0; k < A_COUNT; k++) free(r->a[k].p);
    free(r);
}

/* Stand-alone pieces used by the tests' independent checks. */
float orc_dist_sq(const float *pts, int d, int64_t i, int64_t j) {
    return dist_sq_fma(pts, d, i, j);
}

float orc_fmaf(float a, float b, float c) { return fmaf(a, b, c); }

/*
 * k-th nearest-neighbour squared distance of every point (threshold.py:
 * 133-176: per-row top-k over the distance matrix with the diagonal
 * masked, k-th = largest of the k smallest).  k must lie in [1, n-1].
 * Distances are the filtration's FMA chain.  A sorted insertion into a
 * list of the k smallest values, O(n^2 (d + k)).
 */
int orc_knn_kth(const float *pts, int64_t n, int d, int k, float *out) {
    if (n < 2 || k < 1 || k > n - 1) return -1;
    float *top = (float *)xmalloc(nz(k) * sizeof(float));
    for (int64_t q = 0; q < n; q++) {
        int m = 0; /* entries in top, ascending */
        for (int64_t j = 0; j < n; j++) {
            if (j == q) continue;
            const float v = dist_sq_fma(pts, d, q, j);
            if (m == k && !(v < top[k - 1])) continue;
            int at = m < k ? m : k - 1;
            while (at > 0 && top[at - 1] > v) {
                top[at] = top[at - 1];
                at--;
            }
            top[at] = v;
            if (m < k) m++;
        }
        out[q] = top[k - 1];
    }
    free(top);
    return 0;
}

/* Largest squared distance from each point to any other (threshold.py:
 * 179-200 before its min over points); FMA-chain distances. */
int orc_far_dsq(const float *pts, int64_t n, int d, float *out) {
    for (int64_t q = 0; q < n; q++) {
        float best = 0.0f;
        for (int64_t j = 0; j < n; j++) {
            if (j == q) continue;
            const float v = dist_sq_fma(pts, d, q, j);
            if (v > best) best = v;
        }
        out[q] = best;
    }
    return 0;
}
