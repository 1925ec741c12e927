/*
 * flashph.h -- C ABI of the MI355X (gfx950) exact Vietoris-Rips library
 * (libflashph.so, built from flash-ph_amd/csrc).
 *
 * Conventions
 *   - Every pointer argument is a device pointer unless stated otherwise; the
 *     caller owns all buffers.
 *   - `stream` is a hipStream_t passed as void*; every entry point only
 *     enqueues work on it (stream-ordered, asynchronous, reentrant).
 *   - Entry points taking (ws, ws_bytes) need device workspace: call with
 *     ws == NULL to receive the size in *ws_bytes (nothing is launched), then
 *     call again with a buffer of at least that size.
 *   - Return value: FPH_OK or a negative FPH_ERR_* code.  Nothing throws
 *     across the ABI.
 *   - Data-dependent sizes (edge, triangle, tetrahedron counts) are written to
 *     device memory; the caller reads them and sizes the next stage.
 *
 * Reference citations are relative to the reference's src/flash_ph/.
 */
#ifndef FLASHPH_H
#define FLASHPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FPH_OK 0
#define FPH_ERR_ARG (-1)      /* invalid argument or workspace too small */
#define FPH_ERR_CAPACITY (-2) /* output capacity exceeded (counts still written) */
#define FPH_ERR_LIMIT (-3)    /* size beyond the int32 index range */
#define FPH_ERR_HIP (-4)      /* HIP runtime error */

/* Library version string (host pointer). */
const char *fph_version(void);

/* ---- stage (a): edges ------------------------------------------------- */

/*
 * Screen all pairs i < j of `pts` (n x d, row-major FP32) and append those
 * with dsq < thr_sq, where dsq = fma-chain sum_k (x_i[k]-x_j[k])^2 in FP32,
 * as unordered (key, val) pairs:
 *   key = float_bits(dsq) << 32 | (j(j-1)/2 + i) mod 2^32,  val = j << 32 | i
 * (vals may be NULL: keys only).
 * *count_dev receives the total number of qualifying pairs even when it
 * exceeds `cap` (only the first `cap` are written).
 * Replaces kernels/edge_kernel.py:23-149 (count + emit kernels) and
 * direct_edge_enumerate (kernels/edge_kernel.py:152-224).
 */
int fph_edge_screen(const float *pts, int64_t n, int d, float thr_sq, uint64_t *keys,
                    uint64_t *vals, int64_t cap, uint64_t *count_dev, void *stream);

/*
 * fph_edge_screen for wide points on the matrix cores: Gram products of
 * 128 x 128 point tiles with split-bf16 operands (x = xh + xl) on
 * v_mfma_f32_32x32x16_bf16 with FP32 accumulation screen every pair against
 * t^2 plus a rigorous error margin; candidates are recomputed in the direct
 * FP32 form and kept iff < thr_sq.  Same outputs as fph_edge_screen (same
 * pairs, same dsq bits, unordered); *cand_dev = candidates recomputed
 * (>= edges; the difference is the margin band).  The Gram products are
 * formed on coordinates centred on the cloud's mean, so a common offset
 * does not widen the band.  Workspace: the bf16 splits, norms, centre and
 * its partial sums (about 4 n ceil16(d) + 4 n + 2 KiB d bytes).
 * Replaces kernels/edge_kernel.py:23-224 (the reference computes the direct
 * form for every pair).
 */
int fph_edge_screen_mfma(const float *pts, int64_t n, int d, float thr_sq, uint64_t *keys,
                         uint64_t *vals, int64_t cap, uint64_t *count_dev, uint64_t *cand_dev,
                         void *ws, size_t *ws_bytes, void *stream);

/*
 * Validation mode of fph_edge_screen_mfma: the same screen, which also lists
 * the margin band -- the candidates whose direct FP32 recompute is not below
 * thr_sq -- as val = j << 32 | i in band (capacity band_cap; *band_count =
 * their total, which may exceed band_cap: rerun with more room).  The band
 * and the edges are the pairs fph_pairs_fp64_check re-examines in FP64.
 * No counterpart in the reference (which has no matrix-core screen).
 */
int fph_edge_screen_mfma_band(const float *pts, int64_t n, int d, float thr_sq, uint64_t *keys,
                              uint64_t *vals, int64_t cap, uint64_t *count_dev,
                              uint64_t *cand_dev, uint64_t *band, int64_t band_cap,
                              uint64_t *band_count, void *ws, size_t *ws_bytes, void *stream);

/*
 * FP64 re-examination of m screened pairs (pair_vals: j << 32 | i, as in
 * fph_edge_screen's vals) on the device: per pair the filtration's FP32 FMA
 * chain and the FP64 distance.  stats (device uint64[8], overwritten): [0]
 * pairs, [1] FP32 members (dsq32 < thr_sq), [2] exact members (dsq64 <
 * max_edge_length^2 in FP64), [3] FP32 members that are not exact members,
 * [4] exact members that are not FP32 members, [5] the largest
 * |dsq32 - dsq64| over FP32 members in 1/1024 ulp of dsq32.
 * Used on the margin band and the edges in validation mode; the reference's
 * filtration is defined by the FP32 values (kernels/edge_kernel.py:112-125),
 * so this reports, and does not change, its membership.
 */
int fph_pairs_fp64_check(const float *pts, int64_t n, int d, double max_edge_length,
                         float thr_sq, const uint64_t *pair_vals, int64_t m, uint64_t *stats,
                         void *stream);

/*
 * Sort the E screened pairs by key (ties, possible only for n > 92682, in
 * the reference's emission order) and unpack them into edge_i, edge_j,
 * edge_dsq and edge_filt = sqrt(edge_dsq) (correctly rounded; may be NULL).
 * keys_tmp / vals_tmp: E-element scratch.
 * Replaces _pack_sort_key + argsort(stable=True) + gathers, rips.py:30-50,
 * 188-202.
 */
int fph_edge_sort(const uint64_t *keys_in, const uint64_t *vals_in, int64_t E, int64_t n,
                  uint64_t *keys_tmp, uint64_t *vals_tmp, int32_t *edge_i, int32_t *edge_j,
                  float *edge_dsq, float *edge_filt, void *ws, size_t *ws_bytes, void *stream);

/*
 * fph_edge_sort with ties in dsq broken by the full 64-bit colex index
 * j (j - 1) / 2 + i instead of the reference's 32-bit truncation and
 * emission order: the order (dsq, j, i).  Identical to fph_edge_sort for
 * n <= 92682; beyond, an opt-in departure from the reference's order
 * (rips.py:36-49 notes the truncation) that keeps every diagram's multiset
 * of intervals.
 */
int fph_edge_sort_colex64(const uint64_t *keys_in, const uint64_t *vals_in, int64_t E, int64_t n,
                          uint64_t *keys_tmp, uint64_t *vals_tmp, int32_t *edge_i,
                          int32_t *edge_j, float *edge_dsq, float *edge_filt, void *ws,
                          size_t *ws_bytes, void *stream);

/*
 * Per-vertex CSR of the sorted edges: adj_ptr (n+1), and per half-edge the
 * neighbour (ascending within a row), its edge rank and dsq (2E each).
 * Replaces _build_vertex_csr, rips.py:53-95.
 */
int fph_vertex_csr(const int32_t *edge_i, const int32_t *edge_j, const float *edge_dsq,
                   int64_t E, int64_t n, int32_t *adj_ptr, int32_t *adj_idx,
                   int32_t *adj_rank, float *adj_dsq, void *ws, size_t *ws_bytes,
                   void *stream);

/*
 * COO triplets of the filtration's sparse distance matrix for an external
 * reducer: entries 0..E-1 = (edge_i[e], edge_j[e], sqrt(edge_dsq[e])) in
 * rank order (correctly rounded FP32 sqrt), entries E..E+n-1 = (v, v, 0)
 * for every vertex.  row / col int64, val FP32, E + n entries each.
 * Replaces the COO assembly of _rips_h1h2_giotto, persistence.py:27-45.
 */
int fph_coo_export(const int32_t *edge_i, const int32_t *edge_j, const float *edge_dsq,
                   int64_t E, int64_t n, int64_t *row, int64_t *col, float *val, void *stream);

/* ---- stage (b): minimum spanning forest and H0 ------------------------ */

/* Number of Boruvka rounds issued for n vertices (ceil(log2 n) + 2). */
int fph_boruvka_rounds(int64_t n);
/* Bytes of device scratch fph_boruvka_mst needs for n vertices. */
size_t fph_boruvka_scratch_bytes(int64_t n);

/*
 * Boruvka minimum spanning forest over edge ranks 0..E-1 (the sorted edge
 * order).  mst_mask[e] = 1 for forest edges.  comp_label (n, int64, may be
 * NULL) = smallest vertex id of each vertex's component.  *converged = 1
 * when the issued rounds sufficed (always, for the round count above).
 * Replaces gpu_boruvka_mst, boruvka.py:17-113.
 */
int fph_boruvka_mst(const int32_t *edge_i, const int32_t *edge_j, int64_t E, int64_t n,
                    uint8_t *mst_mask, int64_t *comp_label, int32_t *converged,
                    void *scratch, size_t scratch_bytes, void *stream);

/*
 * fph_boruvka_mst without a scratch argument and without allocating: the
 * round kernels' state lives in a static device array of the library
 * (4 * 4 * FPH_BORUVKA_NOALLOC_MAX_N + 256 bytes = 4 MiB per GPU, held until
 * the process ends; its address is looked up once per device).  Same
 * outputs as fph_boruvka_mst.  FPH_ERR_LIMIT for n > FPH_BORUVKA_NOALLOC_MAX_N
 * (checked before any HIP call).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (two streams, or two host threads)
 * race and must be serialised by the caller; the kernels initialise every
 * word they read, so nothing of a previous call leaks into the next.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).
 */
#define FPH_BORUVKA_NOALLOC_MAX_N 262144
int fph_boruvka_mst_noalloc(const int32_t *edge_i, const int32_t *edge_j, int64_t E, int64_t n,
                            uint8_t *mst_mask, int64_t *comp_label, int32_t *converged,
                            void *stream);

/*
 * H0 diagram: exactly n rows (n x 2 FP32): (0, sqrt(dsq)) for forest edges in
 * rank order, then (0, inf) per component.  mst_idx (int64, capacity n, may
 * be NULL) = forest edge indices ascending; *n_mst (device int64, may be NULL).
 * Replaces persistence.py:169-192.
 */
int fph_h0_diagram(const uint8_t *mst_mask, const float *edge_dsq, int64_t E, int64_t n,
                   float *h0, int64_t *mst_idx, int64_t *n_mst, void *ws, size_t *ws_bytes,
                   void *stream);

/*
 * H0 (n x 2 FP32, the rows of fph_h0_diagram) straight from E screened keys
 * in any order (fph_edge_screen / fph_edge_screen_mfma keys; vals may be
 * NULL there): Boruvka runs on the keys themselves -- unique, and decoded
 * back to (i, j), while n <= 92682 -- with 64-bit minima per component;
 * only the forest (at most n - 1 keys) is sorted.  No edge sort and no CSR,
 * so E may exceed 2^31 (the keys stay in HBM, 8 bytes per edge).
 * *converged as for fph_boruvka_mst.  FPH_ERR_LIMIT for n > 92682.
 * Replaces, for max_dim = 0, rips_filtration + gpu_boruvka_mst + the H0
 * assembly (rips.py:98-218, boruvka.py:17-113, persistence.py:140-186),
 * without the reference's E < 2^31 assertion (kernels/edge_kernel.py:203).
 */
int fph_h0_from_keys(const uint64_t *keys, int64_t E, int64_t n, float *h0, int32_t *converged,
                     void *ws, size_t *ws_bytes, void *stream);

/* ---- stage (d): triangles and H1 cofacet columns ---------------------- */

/*
 * Per edge e = (u, v): tri_off = exclusive scan of the number of triangles
 * whose largest edge is e (E+1 entries); cof_off = exclusive scan of the
 * number of triangles containing e, 0 for forest edges (E+1 entries).
 * totals[0] = T, totals[1] = total cofacet entries (device int64[2]).
 * Replaces _triangle_count_kernel + prefix, kernels/triangle_kernel.py:28-79,
 * 363-385.
 */
int fph_triangle_count(const int32_t *edge_i, const int32_t *edge_j, int64_t E,
                       const int32_t *adj_ptr, const int32_t *adj_idx, const int32_t *adj_rank,
                       const uint8_t *mst_mask, int32_t *tri_off, int32_t *cof_off,
                       int64_t *totals, void *ws, size_t *ws_bytes, void *stream);

/*
 * Triangles in filtration order (global rank E + t): tri_w (third vertex,
 * ascending per owning edge), tri_e (owning = largest edge), tri_ru / tri_rv
 * (ranks of (u,w), (v,w)), tri_cmf (largest non-forest face rank or -1).
 * Replaces _triangle_emit_kernel (kernels/triangle_kernel.py:81-144), the
 * face lookup (reduce_h1_gpu.py:65-100) and the triangle ranking
 * (reduce_h1_gpu.py:509-521).
 */
int fph_triangle_emit(const int32_t *edge_i, const int32_t *edge_j, int64_t E,
                      const int32_t *adj_ptr, const int32_t *adj_idx, const int32_t *adj_rank,
                      const uint8_t *mst_mask, const int32_t *tri_off, int32_t *tri_w,
                      int32_t *tri_e, int32_t *tri_ru, int32_t *tri_rv, int32_t *tri_cmf,
                      void *stream);

/*
 * Cofacet columns of the non-forest edges: ascending triangle global ranks
 * (C = totals[1] entries, rows by cof_off).
 * Replaces _build_cofacet_csr_h1, reduce_h1_gpu.py:107-152.
 */
int fph_cofacets_h1(const int32_t *edge_i, const int32_t *edge_j, int64_t E, int64_t T,
                    const int32_t *adj_ptr, const int32_t *adj_idx, const int32_t *adj_rank,
                    const uint8_t *mst_mask, const int32_t *tri_off, const int32_t *tri_w,
                    const int32_t *cof_off, int64_t C, uint32_t *cof_data, void *ws,
                    size_t *ws_bytes, void *stream);

/* ---- stage (c): apparent pairs ---------------------------------------- */

/*
 * Classify K columns (global rank col_base + c) with cofacet rows
 * (col_off, col_data; global row ranks, row 0 = row_base, nrows rows):
 * kind = 0 skipped (skip_mask), 1 no cofacet (essential), 2 apparent
 * (row_cmf[youngest - row_base] == column rank), 3 residual.  own (nrows,
 * 64-bit) is cleared for fph_reduce, which records the pivot owners in it;
 * pivot[c] = youngest for apparent columns, -1 otherwise; residual columns
 * are listed (unordered) in res_list, their number in *res_count (device).
 * Replaces _gpu_apparent_pair_predetect, reduce_h1_gpu.py:159-223, for H1
 * (col_base 0, row_base E) and H2 (col_base E, row_base E + T).
 */
int fph_apparent(int64_t K, int64_t col_base, const int32_t *col_off, const uint32_t *col_data,
                 const uint8_t *skip_mask, const int32_t *row_cmf, int64_t row_base,
                 int64_t nrows, uint8_t *kind, uint64_t *own, int32_t *pivot, int32_t *res_list,
                 int32_t *res_count, void *stream);

/*
 * Ripser-style apparent-pair scan of K non-forest edges (edge ranks in
 * nonmst_ranks; edges indexed by rank): is_apparent[k] = 1 when a common
 * neighbour w of the edge's endpoints has both connecting edges of lower
 * rank, and pivot_colex[k] = C(c,3) + C(b,2) + a of the triangle with the
 * smallest such w (sorted a < b < c; 0 when not apparent).  max_degree > 0
 * reproduces the reference's GPU loop, which examines at most 2 md steps of
 * the neighbour-list merge (md = max_degree rounded up to a power of two in
 * [64, 2048]); max_degree <= 0 scans completely (the reference's CPU path).
 * The set is a subset of fph_apparent's forest-cleared one.
 * Replaces apparent_pair_scan, kernels/apparent_pair_kernel.py:101-190,
 * 200-313 (not wired into the reference's pipeline).
 */
int fph_apparent_scan(const int32_t *edge_i, const int32_t *edge_j,
                      const int32_t *nonmst_ranks, int64_t K, const int32_t *adj_ptr,
                      const int32_t *adj_idx, const int32_t *adj_rank, int64_t max_degree,
                      int32_t *is_apparent, int64_t *pivot_colex, void *stream);

/*
 * Ripser-style apparent-pair scan for H2: for each of T triangles
 * (tri_v0 < tri_v1 < tri_v2, largest edge rank tri_max_rank), the youngest
 * tetrahedron cofacet by (largest edge rank, colex4) over the common
 * neighbours of its three vertices; is_apparent = 1 when that rank equals
 * the triangle's and the triangle is the tetrahedron's oldest face (the
 * added vertex is below the vertex off the largest edge), with pivot_colex
 * = the tetrahedron's colex index C(v3,4)+C(v2,3)+C(v1,2)+v0 (else 0).
 * max_degree > 0 bounds the walk as the reference's GPU loop does (md =
 * max_degree rounded up to a power of two in [64, 2048], 3 md merge steps);
 * <= 0: no bound (the reference's Numba path).  FPH_ERR_LIMIT past 2^33
 * triangles.
 * Replaces h2_apparent_pair_scan, kernels/apparent_pair_h2_kernel.py:
 * 61-153, 161-317, 328-441 (not wired into the reference's pipeline).
 */
int fph_apparent_scan_h2(const int32_t *tri_v0, const int32_t *tri_v1, const int32_t *tri_v2,
                         const int32_t *tri_max_rank, int64_t T, const int32_t *adj_ptr,
                         const int32_t *adj_idx, const int32_t *adj_rank, int64_t max_degree,
                         int32_t *is_apparent, int64_t *pivot_colex, void *stream);

/* ---- stage (e): residual reduction ------------------------------------ */

/* Bytes of device scratch fph_reduce needs for K columns. */
size_t fph_reduce_scratch_bytes(int64_t K);

/*
 * Z/2 reduction of the residual columns listed by fph_apparent.  On return
 * pivot[c] holds each residual column's final pivot row (global rank) or -1
 * (reduced to zero: essential), and own[r - row_base] = (owner + 1) << 32 |
 * pool offset for every pivot row.  Columns are written to `pool`
 * (pool_cap 32-bit words, as [length, rows...]; the buffer must extend 64
 * words past pool_cap).  status (device int32[8], 8-byte aligned): [0] = 0
 * ok, 1 pool too small (retry with a pool of more than 32 * status[5]
 * words), 3 inconsistent input, 5 a waiting wavefront of the second phase
 * saw no progress for wait_seconds (<= 0: 120 s); [1] = 1 once the second
 * phase ran; [2] = column additions performed; [3] = longest reduced
 * column; [4] = additions too long for on-chip memory; [5] = pool words
 * handed out, in 32-word lines; [6..7] = uint64 entries merged (sum over
 * additions of both input lengths and the result length).
 * seg_cols (nseg+1 column offsets, device, may be NULL when nseg == 1)
 * splits the columns into independent segments (the point clouds of a
 * batch: no row is shared between segments); one workgroup reduces each.
 * Replaces _general_cohomology_reduce_v2 (cohomology_general.py:202-371)
 * with its pool prefill (_numba_utils.py:68-104, reduce_h1_gpu.py:242-303).
 */
int fph_reduce(int64_t K, const int32_t *col_off, const uint32_t *col_data, int64_t row_base,
               int64_t nrows, const int32_t *seg_cols, int64_t nseg, const uint8_t *kind,
               uint64_t *own, int32_t *pivot, const int32_t *res_list,
               const int32_t *res_count, uint32_t *pool, int64_t pool_cap, void *scratch,
               size_t scratch_bytes, int32_t *status, int wait_seconds, void *stream);

/*
 * Replay of the reference's own reduction after fph_reduce (one segment):
 * every residual column (res_list, descending index) is rebuilt by adding,
 * at each pivot, the column of its final owner from fph_reduce (own[],
 * pivot[]) -- the CSR column of an apparent owner, else the owner's rebuilt
 * column -- until it reaches its final pivot.  While every reduced column
 * fits the reference's store and no working column passes MAX_COL, these
 * are exactly the reference's additions (standard reduction, descending
 * order), so the reported totals decide whether fph_reduce's pairs are the
 * reference's.  Needs ceil(nrows / 32) <= 2496 (FPH_ERR_LIMIT otherwise).
 * pool: pool_cap words plus 64 of slack, scratch: fph_reduce_scratch_bytes.
 * status (device int32[8], 8-byte aligned): [0] = 0 ok, 1 pool too small
 * (the rebuilt columns need more than pool_cap words), 3 inconsistent
 * owners, 5 no progress for wait_seconds (<= 0: 120 s); [1] = 1 once run;
 * [2] = column additions; [3] = longest working column (initial columns
 * included); [4..5] = uint64 entries the reference's store holds (apparent
 * columns plus non-zero reduced columns); [6] = columns it holds.
 * Follows _general_cohomology_reduce_v2 (cohomology_general.py:255-366),
 * the store test (reduce_h1_gpu.py:283-289, cohomology_general.py:337-345)
 * and the MAX_COL cut (_numba_utils.py:10, 44-63).
 */
int fph_reduce_replay(int64_t K, const int32_t *col_off, const uint32_t *col_data,
                      int64_t row_base, int64_t nrows, const uint8_t *kind, const uint64_t *own,
                      const int32_t *pivot, const int32_t *res_list, const int32_t *res_count,
                      uint32_t *pool, int64_t pool_cap, void *scratch, size_t scratch_bytes,
                      int32_t *status, int wait_seconds, void *stream);

/*
 * The reference's own reduction semantics, for results identical to it in
 * every case: per segment, apparent columns enter a store of seg_pool[2s+1]
 * entries in ascending column order, then residual columns are reduced one
 * at a time in descending rank order; a column that does not fit in the
 * store is not stored (later columns reaching its pivot stop there) and
 * column additions keep at most 1048576 rows.  Where the store never
 * fills, the pairs equal fph_reduce's.  seg_pool (device int64, 5 per
 * segment): [0] word offset in `pool` of the segment's store (at least
 * capacity + segment columns + 64 words, beyond the apparent region of
 * col_off[K] + K words), [1] its capacity in reference entries, [2] global
 * rank of the segment's first row, [3] its number of rows, [4] word offset
 * of a zeroed bitmap of ceil(rows / 32) words in `pool` for the working
 * column, or -1 when ceil(rows/32) + ceil(rows/1024) <= 32768 (kept in LDS).
 * pool_start: first pool word after the stores and bitmaps.  variant bit 0:
 * 0 = working column as a bitmap (status[0] = 4 when a column would exceed
 * max_col rows: rerun with bit 0 set), 1 = as a sorted list.  Bit 1 (with
 * seg_pool[0] = -1, capacity ignored): the complete reduction instead --
 * every column stored (bump-allocated in the pool), no truncation; the
 * pairs of fph_reduce, in the reference's serial order (faster than
 * fph_reduce on dense complexes with long columns).  A bitmap in `pool`
 * has a window of rows in LDS (the rows from the current pivot on, about
 * 9.5e5 of them); bit 3 turns the window off, bits 8..15 cap it at that
 * many units of 32768 rows (tests; 0 = as many as LDS holds).  max_col: rows kept by
 * a column addition in the reference semantics, 0 = the reference's
 * MAX_COL of 1048576 (a smaller value reproduces the truncation branch on
 * small complexes, for tests).  status as for fph_reduce.
 * Replaces _general_cohomology_reduce_v2 (cohomology_general.py:202-371)
 * with the pool sizing of reduce_h1_gpu.py:282-289 and the prefill of
 * _numba_utils.py:68-104, including their capacity limits.
 */
int fph_reduce_serial(int64_t K, const int32_t *col_off, const uint32_t *col_data,
                      int64_t row_base, int64_t nrows, const int32_t *seg_cols, int64_t nseg,
                      const int64_t *seg_pool, int64_t pool_start, const uint8_t *kind,
                      uint64_t *own, int32_t *pivot, const int32_t *res_list,
                      const int32_t *res_count, uint32_t *pool, int64_t pool_cap,
                      void *scratch, size_t scratch_bytes, int32_t *status, int variant,
                      int64_t max_col, void *stream);

/*
 * fph_reduce_serial's results (same arguments, same pairs, store contents,
 * additions and status codes) with many columns in flight: one wavefront
 * per column takes the additions the stored columns allow and the columns
 * settle in the reference's order (a column stops at a pivot, and is stored
 * or dropped, only after every earlier column has settled).  Segments run
 * concurrently, most residual columns first, each on one XCD.  seg_pool[4]
 * is ignored: working bitmaps are per wavefront, in LDS when ceil(max_rows
 * / 32) + ceil(max_rows / 1024) <= 20224 words, else a window of rows in
 * LDS over a slot in `ws` (bit 3 of variant: the whole bitmap in the slot,
 * summaries in LDS; bits 8..15: at most that many 1024-word window units,
 * for tests; FPH_ERR_LIMIT past 20224 summary words, about 2.1e7 rows;
 * bits 16..23: at most that many unsettled columns per segment, default
 * a quarter of an XCD's wavefronts for bitmaps in LDS, an eighth else).
 * max_rows: the largest seg_pool[3].  Workspace by size query (ws == NULL).
 * variant bit 0 (the sorted-list form) is not taken: status[0] = 4 asks the
 * caller for fph_reduce_serial with bit 0; bit 2 puts every working bitmap
 * in `ws` (for tests of that form on small clouds).  Candidates are staged
 * in `pool` from pool_start on (status 1 when it is too small).
 * No wavefront waits for another; wait_seconds (<= 0: 120 s) bounds how
 * long one may find no work before the call gives up with status[0] = 5,
 * which is also set if a position was left unsettled (protocol checks).
 * Replaces _general_cohomology_reduce_v2 (cohomology_general.py:202-371)
 * like fph_reduce_serial.
 */
int fph_reduce_pipe(int64_t K, const int32_t *col_off, const uint32_t *col_data,
                    int64_t row_base, int64_t nrows, const int32_t *seg_cols, int64_t nseg,
                    const int64_t *seg_pool, int64_t max_rows, int64_t pool_start,
                    const uint8_t *kind, uint64_t *own, int32_t *pivot, const int32_t *res_list,
                    const int32_t *res_count, uint32_t *pool, int64_t pool_cap, void *scratch,
                    size_t scratch_bytes, void *ws, size_t *ws_bytes, int32_t *status,
                    int variant, int64_t max_col, int wait_seconds, void *stream);

/* ---- diagrams --------------------------------------------------------- */

/*
 * Rows (birth, death) of a dimension in the reference's order: apparent
 * pairs ascending, reduced pairs descending, essentials descending (or
 * ascending when ess_ascending), then stably by birth value; only pairs with
 * death > birth.  Birth of column c = edge_filt[col_edge ? col_edge[c] : c];
 * death of row r = edge_filt[row_edge[r - row_base]].  out: K x 2 FP32
 * capacity.  Segments (seg_cols: nseg+1 column offsets, seg_edges: nseg+1
 * edge offsets, seg_ess_asc: per-segment flag; all NULL for one segment
 * spanning [0, K) and edges [0, E)) are the clouds of a batch: segment s
 * writes counts[s] rows at out[seg_cols[s] ...].  out_edges (K x 2 int32,
 * may be NULL): per row the edge whose length is the birth and the one
 * whose length is the death (-1 for an infinite death) -- the diameter
 * edges of the critical simplices, for gradients.
 * Replaces reduce_h1_gpu.py:495-502, 574-604 (and persistence.py:202-205
 * for H2).
 */
int fph_diagram(int64_t K, const uint8_t *kind, const int32_t *col_off,
                const uint32_t *col_data, const int32_t *pivot, const int32_t *col_edge,
                int64_t row_base, const int32_t *row_edge, const float *edge_filt, int64_t E,
                int ess_ascending, const int32_t *seg_cols, const int32_t *seg_edges,
                const uint8_t *seg_ess_asc, int64_t nseg, float *out, int32_t *out_edges,
                int32_t *counts, void *ws, size_t *ws_bytes, void *stream);

/* ---- H2: clearing cascade, tetrahedra, cofacet columns ---------------- */

/*
 * cleared[t] = 1 for every triangle that is the pivot of an H1 column
 * (apparent or reduced, including zero persistence).  nrows = T.
 * Replaces the all_pivots cascade, cohomology_general.py:171-173.
 */
int fph_clear_pivots(int64_t K, const uint8_t *kind, const int32_t *pivot, int64_t row_base,
                     int64_t nrows, uint8_t *cleared, void *stream);

/*
 * tet_off = exclusive scan of the number of tetrahedra whose largest edge is
 * e (E+1 entries); *total = Q (device int64).
 * Replaces _tet_count_kernel + prefix, kernels/triangle_kernel.py:152-204,
 * 468-489.
 */
int fph_tet_count(int64_t E, const int32_t *adj_ptr, const int32_t *adj_idx,
                  const int32_t *adj_rank, const int32_t *tri_off, const int32_t *tri_w,
                  int32_t *tet_off, int64_t *total, void *ws, size_t *ws_bytes, void *stream);

/*
 * Tetrahedra in (max_rank, colex4) order (global rank E + T + q):
 * tet_key = y << 32 | x for the two vertices off the largest edge (any n;
 * the reference packs 16-bit fields and so needs n < 65536,
 * persistence.py:125-129), tet_e = largest edge, tet_cmf = largest
 * uncleared triangle face (global rank) or -1.
 * Replaces _tet_emit_kernel (kernels/triangle_kernel.py:206-270) and the
 * colex4 order of kernels/apparent_pair_h2_kernel.py:52-58.
 */
int fph_tet_emit(int64_t E, const int32_t *edge_i, const int32_t *edge_j,
                 const int32_t *adj_ptr, const int32_t *adj_idx, const int32_t *adj_rank,
                 const int32_t *tri_off, const int32_t *tri_w, const int32_t *tri_ru,
                 const int32_t *tri_rv, const uint8_t *cleared, const int32_t *tet_off,
                 uint64_t *tet_key, int32_t *tet_e, int32_t *tet_cmf, void *stream);

/*
 * cof_off (T+1) = exclusive scan of the number of tetrahedra containing each
 * uncleared triangle; *total = total entries (device int64).
 * Replaces the 3-way merge count of kernels/apparent_pair_h2_kernel.py:
 * 202-260 / _numba_utils.py:186-230.
 */
int fph_cofacets_h2_count(int64_t T, const int32_t *edge_i, const int32_t *edge_j,
                          const int32_t *adj_ptr, const int32_t *adj_idx, const int32_t *tri_w,
                          const int32_t *tri_e, const uint8_t *cleared, int32_t *cof_off,
                          int64_t *total, void *ws, size_t *ws_bytes, void *stream);

/*
 * Cofacet columns of the uncleared triangles: ascending tetrahedron global
 * ranks.  Restates the H2 column layout of cohomology_general.py:32-199
 * (columns = triangles, cofacets = tetrahedra).
 */
int fph_cofacets_h2(int64_t T, int64_t E, int64_t Q, const int32_t *edge_i,
                    const int32_t *edge_j, const int32_t *adj_ptr, const int32_t *adj_idx,
                    const int32_t *adj_rank, const int32_t *tri_w, const int32_t *tri_e,
                    const uint8_t *cleared, const int32_t *tet_off, const uint64_t *tet_key,
                    const int32_t *cof_off, int64_t C, uint32_t *cof_data, void *ws,
                    size_t *ws_bytes, void *stream);

/* ---- batches of point clouds ------------------------------------------ */
/*
 * A batch is C clouds concatenated: cloud c has n_c = vbase[c+1] - vbase[c]
 * points of dimension dims[c], stored row-major at pts + foff[c] (floats).
 * Vertices get global ids vbase[c] + local id.  A batch is one Vietoris-Rips
 * complex (no edges between clouds); after fph_edge_sort_batch the
 * single-cloud entry points (CSR, MST, triangles, cofacets, apparent pairs)
 * run on it unchanged, and fph_reduce / fph_diagram take the per-cloud
 * column segments.  The reference runs clouds one by one through the same
 * functions; these entry points have no other counterpart.
 */

#define FPH_KNN_MAX_K 128

/*
 * Per-cloud threshold of auto_threshold (threshold.py:20-78): the k-th
 * nearest-neighbour distance of every point (k clamped to [1, n_c - 1]),
 * then the linear-interpolation quantile at percentile/100 (clamped to
 * [0, 100]) in FP32 as torch.quantile; t_out[c] = 0 for n_c <= 1.
 * Distances are the filtration's FP32 FMA-chain values, correctly rounded
 * sqrt.  Query tiles of 256 points: qtile_off (C+1, exclusive scan of
 * ceil(n_c / 256)); max_dim_host = max_c dims[c].  Optional outputs:
 * kth_dist (N floats), thr_sq_out[c] = fp32(double(t)^2).  k <= 128 keeps
 * the k smallest distances in registers; larger k selects the k-th one by
 * radix over the FP32 bits (four passes over the cloud per query).
 * Replaces _tiled_knn (threshold.py:133-176) and the quantile at :77.
 */
int fph_knn_threshold_batch(const float *pts, const int64_t *foff, const int32_t *dims,
                            const int64_t *vbase, const int64_t *qtile_off, int64_t n_qtiles,
                            int64_t C, int64_t N, int k, double percentile, int max_dim_host,
                            float *kth_dist, float *t_out, float *thr_sq_out, void *ws,
                            size_t *ws_bytes, void *stream);

/*
 * Per-cloud enclosing radius min_x max_y |x - y| (threshold.py:81-116,
 * 179-200), with the filtration's FP32 distances; 0 for n_c <= 1.  Same
 * layout arguments as fph_knn_threshold_batch.
 */
int fph_enclosing_radius_batch(const float *pts, const int64_t *foff, const int32_t *dims,
                               const int64_t *vbase, const int64_t *qtile_off,
                               int64_t n_qtiles, int64_t C, int max_dim_host, float *r_out,
                               void *ws, size_t *ws_bytes, void *stream);

/*
 * Screen every cloud against its own FP32 threshold thr_sq[c] (strict <),
 * with the FMA chain of fph_edge_screen.  Tiles of 64 x 64 points:
 * tile_off (C+1) = exclusive scan of nt_c (nt_c + 1) / 2 with
 * nt_c = ceil(n_c / 64).  keys = bits(dsq) << 32 | local colex, vals =
 * global j << 32 | global i; *count_dev = total (may exceed cap: rerun).
 * Replaces direct_edge_enumerate (kernels/edge_kernel.py:152-224) per cloud.
 */
int fph_edge_screen_batch(const float *pts, const int64_t *foff, const int32_t *dims,
                          const int64_t *vbase, const float *thr_sq, const int64_t *tile_off,
                          int64_t n_tiles, int64_t C, uint64_t *keys, uint64_t *vals,
                          int64_t cap, uint64_t *count_dev, void *stream);

/*
 * Edges grouped by cloud, each cloud in its own filtration order (the
 * reference's stable key sort, rips.py:188-202, per cloud; requires every
 * n_c <= 92682 so that keys are unique).  edge_off (C+1): the edge range of
 * each cloud.  edge_filt = sqrt(dsq) correctly rounded.
 */
int fph_edge_sort_batch(const uint64_t *keys_in, const uint64_t *vals_in, int64_t E,
                        const int64_t *vbase, int64_t C, int32_t *edge_i, int32_t *edge_j,
                        float *edge_dsq, float *edge_filt, int32_t *edge_off, void *ws,
                        size_t *ws_bytes, void *stream);

/*
 * H0 of every cloud at rows vbase[c] ... vbase[c+1] - 1 of h0 (N x 2):
 * forest edges of the cloud in filtration order as (0, sqrt(dsq)), then
 * (0, inf) per component (persistence.py:169-192, per cloud).
 */
int fph_h0_diagram_batch(const uint8_t *mst_mask, const float *edge_dsq, int64_t E,
                         const int32_t *edge_off, const int64_t *vbase, int64_t C, int64_t N,
                         float *h0, void *ws, size_t *ws_bytes, void *stream);

/* ---- gradients -------------------------------------------------------- */

/*
 * Gradient of a loss on diagram entries with respect to the points (same
 * batch layout as fph_edge_screen_batch; one cloud: C = 1, foff = {0, n d},
 * vbase = {0, n}).  Entry m has upstream gradient contrib_val[m] and sits on
 * edge contrib_edge[m] (-1: none, e.g. an H0 birth or an infinite death),
 * as given by fph_diagram's out_edges and the H0 forest edges.  With
 * w_e = (sum of the entries on e) / edge_filt[e]:
 *   grad_pts[v] = sum over the CSR edges e = (v, u) of w_e (x_v - x_u).
 * Sums run in a fixed order (deterministic).  grad_pts has the layout of
 * pts.  No counterpart in the reference (which has no gradient path).
 */
int fph_backward_scatter(const float *pts, const int64_t *foff, const int32_t *dims,
                         const int64_t *vbase, int64_t C, int64_t N, const float *edge_filt,
                         int64_t E, const int32_t *adj_ptr, const int32_t *adj_idx,
                         const int32_t *adj_rank, const int32_t *contrib_edge,
                         const float *contrib_val, int64_t M, float *grad_pts, void *ws,
                         size_t *ws_bytes, void *stream);

/* ---- graph-captured H0 pipeline ---------------------------------------- */

/*
 * Stages (a) and (b) in capacity form, for graph capture: screen (screen 0
 * direct, 1 split-bf16) -> sort -> vertex CSR -> Boruvka -> H0 diagram with
 * every size fixed at `cap` edge slots on the host.  The live edge count is
 * count_dev[0] (device uint64[2]; [1] = split-bf16 candidates); slots past
 * it hold edge_i = edge_j = -1 and adj_idx = adj_rank = -1 beyond
 * adj_ptr[n].  Requires 2 <= n <= 92682.  Results equal fph_edge_screen +
 * fph_edge_sort + fph_vertex_csr + fph_boruvka_mst + fph_h0_diagram
 * whenever count_dev[0] <= cap; otherwise the edges past cap are lost and
 * the caller retries with cap >= count_dev[0].  Buffers: edge_* and
 * mst_mask cap entries, adj_idx/rank/dsq 2 cap, adj_ptr n+1, h0 2n floats,
 * converged int32[1].  Records kernels only: no memset node, no host
 * synchronisation.  Replaces, for repeated calls on one shape, the
 * host-driven sequence of persistence.py:140-186.
 */
int fph_h0_pipeline(const float *pts, int64_t n, int d, float thr_sq, int screen, int64_t cap,
                    uint64_t *count_dev, int32_t *edge_i, int32_t *edge_j, float *edge_dsq,
                    float *edge_filt, int32_t *adj_ptr, int32_t *adj_idx, int32_t *adj_rank,
                    float *adj_dsq, uint8_t *mst_mask, float *h0, int32_t *converged, void *ws,
                    size_t *ws_bytes, void *stream);

/*
 * Captures fph_h0_pipeline on these buffers (ws from its size query) into a
 * hipGraph and instantiates it; *graph_out receives an opaque handle.  Every
 * buffer must outlive the graph.  fph_graph_launch replays it on `stream`
 * (refill pts first); fph_graph_destroy frees it.  No counterpart in the
 * reference (which launches every stage from Python each call).
 */
int fph_h0_graph_create(const float *pts, int64_t n, int d, float thr_sq, int screen,
                        int64_t cap, uint64_t *count_dev, int32_t *edge_i, int32_t *edge_j,
                        float *edge_dsq, float *edge_filt, int32_t *adj_ptr, int32_t *adj_idx,
                        int32_t *adj_rank, float *adj_dsq, uint8_t *mst_mask, float *h0,
                        int32_t *converged, void *ws, size_t ws_bytes, void **graph_out);
int fph_graph_launch(void *graph, void *stream);
int fph_graph_destroy(void *graph);

/* ---- autotune table ---------------------------------------------------- */

/*
 * Launch parameters that change the tiling, never the results.  Values are
 * kept per bucket (floor(log2 size), floor(log2 d)), host memory only; 0 =
 * the default.  Keys: SCREEN_TILE by (n, d) of fph_edge_screen;
 * ENUM_BLOCK by (E, 0) for the link kernels of triangles, tetrahedra and H1
 * cofacets (one wavefront per edge; threads per workgroup = 64 x edges per
 * workgroup)
 * (by (T, 0) for fph_cofacets_h2_count and fph_cofacets_h2); APPARENT_BLOCK by
 * (K, 0) for fph_apparent and fph_clear_pivots; QUEUE_WGS by (K, 0) for
 * fph_reduce with nseg == 1.  The reference picks its
 * Triton BLOCK sizes as constants (kernels/edge_kernel.py:193-222,
 * kernels/triangle_kernel.py:366-397, reduce_h1_gpu.py:159-223); this table
 * replaces them with values a tuner measures (flash_ph.autotune).
 */
#define FPH_TUNE_SCREEN_TILE 0    /* points per tile side: 32, 64, 128 (default) */
#define FPH_TUNE_ENUM_BLOCK 1     /* threads per workgroup: 64 .. 1024 (default 256) */
#define FPH_TUNE_APPARENT_BLOCK 2 /* threads per workgroup: 64 .. 1024 (default 256) */
#define FPH_TUNE_QUEUE_WGS 3      /* fph_reduce queue workgroups, one segment: 1 .. 64 (default 16 from 8192 columns, else 1) */
#define FPH_TUNE_NKNOBS 4

/* Candidate values of `knob` into out[0..cap); returns their number. */
int fph_autotune_candidates(int knob, int32_t *out, int cap);
/* Sets the value for the bucket of (size, d); value 0 restores the default.
 * FPH_ERR_ARG for an unknown knob or a value not among the candidates. */
int fph_autotune_set(int knob, int64_t size, int64_t d, int value);
/* Value in effect for (size, d), or FPH_ERR_ARG. */
int fph_autotune_query(int knob, int64_t size, int64_t d);
/* Clears every bucket back to the defaults. */
void fph_autotune_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* FLASHPH_H */
