// Stage (a): Vietoris-Rips edge enumeration, filtration sort and vertex CSR.
//
// Replaces kernels/edge_kernel.py:23-224 (count + emit over B x B tiles),
// rips.py:30-50,188-202 (packed key + stable argsort) and rips.py:53-95
// (per-vertex CSR) of the reference.
#include "fph_common.h"

#include <rocprim/rocprim.hpp>

namespace {

constexpr int kChunk = 32;    // dimensions staged in LDS per pass
constexpr int kThreads = 256; // 16 x 16 threads

// Tile index -> (I, J) with I <= J, tiles ordered by J then I.
__device__ __forceinline__ void tile_coords(long long t, int* I, int* J) {
    long long j = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= t) j++;
    while (j * (j + 1) / 2 > t) j--;
    *J = (int)j;
    *I = (int)(t - j * (j + 1) / 2);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Selection of a thread's R x R block (rows i0 + ty*R + r, columns
// j0 + tx*R + c): slot (r, c) is one compare whose ballot gives the wave's
// kept lanes.  CHECKED adds the index tests of tiles that touch the
// diagonal or the end of the cloud.
template <int R, bool CHECKED>
struct ScreenSel {
    const float (&acc)[R][R];
    int i0, j0, n, I, J;
    float thr_sq;

    __device__ __forceinline__ bool kept(int r, int c) const {
        const bool in = acc[r][c] < thr_sq;
        if constexpr (!CHECKED) {
            return in;
        } else {
            const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
            const int i = i0 + ty * R + r, j = j0 + tx * R + c;
            return in && i < n && j < n && (I < J || i < j);
        }
    }

    // The wave's kept pairs: a sum of ballot counts (scalar).
    __device__ __forceinline__ int count() const {
        int total = 0;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int c = 0; c < R; c++) total += __popcll(__ballot(kept(r, c)));
        return total;
    }

    // Writes the kept pairs from slot `pos` on, in (r, c, lane) order;
    // slots whose ballot is zero cost one compare.
    __device__ __forceinline__ void write(unsigned long long pos,
                                          unsigned long long* __restrict__ keys,
                                          unsigned long long* __restrict__ vals,
                                          long long cap) const {
        const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
        const unsigned long long lt = fph_lanemask_lt();
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int c = 0; c < R; c++) {
                const bool ok = kept(r, c);
                const unsigned long long b = __ballot(ok);
                if (b == 0ull) continue;
                if (ok) {
                    const unsigned long long at = pos + __popcll(b & lt);
                    if (at < (unsigned long long)cap) {
                        const unsigned long long i = (unsigned long long)(i0 + ty * R + r);
                        const unsigned long long j = (unsigned long long)(j0 + tx * R + c);
                        unsigned bits;
                        memcpy(&bits, &acc[r][c], 4);
                        const unsigned long long colex = j * (j - 1) / 2 + i;
                        keys[at] = ((unsigned long long)bits << 32) | (colex & 0xFFFFFFFFull);
                        if (vals) vals[at] = (j << 32) | i;
                    }
                }
                pos += __popcll(b);
            }
    }
};

// One workgroup per upper-triangular TILE x TILE tile (TILE = 32, 64 or 128,
// an autotuned launch parameter).  Each thread owns an R x R block of pairs
// (rows ty*R+r, columns tx*R+c, R = TILE/16) and accumulates
//     acc = fma(x_i[k] - x_j[k], x_i[k] - x_j[k], acc),   k = 0 .. d-1
// in the reference's order (edge_kernel.py:112-123 with fused multiply-add),
// so the values do not depend on TILE.  Pairs with acc < thr_sq (strict,
// FP32) are appended, unordered, as
//     key = float_bits(acc) << 32 | (j(j-1)/2 + i) mod 2^32     (rips.py:41-49)
//     val = j << 32 | i
// through one atomic per workgroup.  The count keeps growing past `cap` so the
// caller can size a retry.
template <int TILE>
__global__ __launch_bounds__(kThreads) void k_edge_screen(
    const float* __restrict__ pts, int n, int d, float thr_sq, long long tile0,
    unsigned long long* __restrict__ keys, unsigned long long* __restrict__ vals,
    long long cap, unsigned long long* __restrict__ count) {
    constexpr int R = TILE / 16;
    __shared__ __align__(16) float fa[kChunk * TILE];
    __shared__ __align__(16) float fb[kChunk * TILE];
    int I, J;
    tile_coords(tile0 + (long long)blockIdx.x, &I, &J);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = I * TILE, j0 = J * TILE;

    float acc[R][R];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int c = 0; c < R; c++) acc[r][c] = 0.0f;

    for (int k0 = 0; k0 < d; k0 += kChunk) {
        const int kc = min(kChunk, d - k0);
        __syncthreads();
        // one thread per (point, side): the point's kc coordinates go to
        // column p of fa or fb (consecutive threads, consecutive columns)
        for (int idx = threadIdx.x; idx < 2 * TILE; idx += kThreads) {
            const int side = idx / TILE, p = idx - side * TILE;
            const int g = (side ? j0 : i0) + p;
            float* dst = (side ? fb : fa) + p;
            const float* src = pts + (long long)g * d + k0;
            if (g < n) {
                // unrolled so that 8 coordinates' loads are in flight
                // before the first LDS write waits for its data
#pragma unroll 8
                for (int k = 0; k < kc; k++) dst[k * TILE] = src[k];
            } else {
                for (int k = 0; k < kc; k++) dst[k * TILE] = 0.0f;
            }
        }
        __syncthreads();
        for (int k = 0; k < kc; k++) {
            float av[R], bv[R];
#pragma unroll
            for (int r = 0; r < R; r++) av[r] = fa[k * TILE + ty * R + r];
#pragma unroll
            for (int c = 0; c < R; c++) bv[c] = fb[k * TILE + tx * R + c];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int c = 0; c < R; c++) {
                    const float diff = av[r] - bv[c];
                    acc[r][c] = __builtin_fmaf(diff, diff, acc[r][c]);
                }
        }
    }

    // Tiles strictly above the diagonal and inside the cloud need no index
    // tests (I < J and j0 + TILE <= n give i < i0 + TILE <= j0 <= j < n).
    if (I < J && j0 + TILE <= n) {
        const ScreenSel<R, false> sel{acc, i0, j0, n, I, J, thr_sq};
        const int total = sel.count();
        const unsigned long long pos = fph_wg_reserve<kThreads>(total, count);
        if (total) sel.write(pos, keys, vals, cap);
    } else {
        const ScreenSel<R, true> sel{acc, i0, j0, n, I, J, thr_sq};
        const int total = sel.count();
        const unsigned long long pos = fph_wg_reserve<kThreads>(total, count);
        if (total) sel.write(pos, keys, vals, cap);
    }
}

// Reference emission order of edge (i, j): tile (i/32, j/32) row-major over
// the 32-point block grid, then (i mod 32, j mod 32) row-major
// (edge_kernel.py:133-149, 207-210).  Only consulted when the packed key can
// tie (n > 92682, where j(j-1)/2 + i no longer fits in 32 bits).
__global__ void k_emit_order(const unsigned long long* __restrict__ vals, long long E,
                             long long nb32, unsigned long long* __restrict__ ord,
                             unsigned* __restrict__ idx) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const unsigned long long v = vals[e];
    const unsigned long long i = v & 0xFFFFFFFFull, j = v >> 32;
    ord[e] = (((i >> 5) * (unsigned long long)nb32 + (j >> 5)) << 10) | ((i & 31) << 5) | (j & 31);
    idx[e] = (unsigned)e;
}

// Tie key of the exact order: the full colex index j (j - 1) / 2 + i.
__global__ void k_colex_order(const unsigned long long* __restrict__ vals, long long E,
                              unsigned long long* __restrict__ ord, unsigned* __restrict__ idx) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const unsigned long long v = vals[e];
    const unsigned long long i = v & 0xFFFFFFFFull, j = v >> 32;
    ord[e] = j * (j - 1) / 2 + i;
    idx[e] = (unsigned)e;
}

// Exact order: the key's low half becomes the edge's position in the colex
// pre-order, so one full-width sort gives (dsq, colex) with unique keys.
__global__ void k_rank_low(unsigned long long* __restrict__ keys, long long E) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    keys[e] = (keys[e] & 0xFFFFFFFF00000000ull) | (unsigned long long)e;
}

__global__ void k_gather_kv(const unsigned long long* __restrict__ keys,
                            const unsigned long long* __restrict__ vals,
                            const unsigned* __restrict__ idx, long long E,
                            unsigned long long* __restrict__ ko,
                            unsigned long long* __restrict__ vo) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const unsigned s = idx[e];
    ko[e] = keys[s];
    vo[e] = vals[s];
}

__global__ void k_edge_unpack(const unsigned long long* __restrict__ keys,
                              const unsigned long long* __restrict__ vals, long long E,
                              int* __restrict__ ei, int* __restrict__ ej,
                              float* __restrict__ dsq, float* __restrict__ filt) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const unsigned long long v = vals[e];
    const unsigned bits = (unsigned)(keys[e] >> 32);
    float f;
    memcpy(&f, &bits, 4);
    ei[e] = (int)(v & 0xFFFFFFFFull);
    ej[e] = (int)(v >> 32);
    dsq[e] = f;
    if (filt) filt[e] = fph_sqrt_rn(f);
}

// Half-edges (i, j) and (j, i) of every edge; an unused slot of a capacity-
// sized edge list (ei = -1) gives two all-ones keys, which sort last.
__global__ void k_half_edges(const int* __restrict__ ei, const int* __restrict__ ej, long long E,
                             unsigned long long* __restrict__ hk, unsigned* __restrict__ hv) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    if (ei[e] < 0) {
        hk[e] = hk[E + e] = ~0ull;
        hv[e] = hv[E + e] = 0xFFFFFFFFu;
        return;
    }
    const unsigned long long i = (unsigned)ei[e], j = (unsigned)ej[e];
    hk[e] = (i << 32) | j;
    hv[e] = (unsigned)e;
    hk[E + e] = (j << 32) | i;
    hv[E + e] = (unsigned)e;
}

__global__ void k_csr_unpack(const unsigned long long* __restrict__ hk,
                             const unsigned* __restrict__ hv, long long H,
                             const float* __restrict__ edge_dsq, int* __restrict__ adj_idx,
                             int* __restrict__ adj_rank, float* __restrict__ adj_dsq) {
    long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    if (hk[h] == ~0ull) {  // unused slot (capacity-sized lists), past the last row
        adj_idx[h] = -1;
        adj_rank[h] = -1;
        adj_dsq[h] = 0.0f;
        return;
    }
    const unsigned r = hv[h];
    adj_idx[h] = (int)(hk[h] & 0xFFFFFFFFull);
    adj_rank[h] = (int)r;
    adj_dsq[h] = edge_dsq[r];
}

__global__ void k_csr_ptr(const unsigned long long* __restrict__ hk, long long H, long long n,
                          int* __restrict__ ptr) {
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    const unsigned long long key = (unsigned long long)v << 32;
    long long lo = 0, hi = H;
    while (lo < hi) {
        long long mid = (lo + hi) >> 1;
        if (hk[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    ptr[v] = (int)lo;
}

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

int bit_width_u64(unsigned long long x) {
    int b = 0;
    while (x) { b++; x >>= 1; }
    return b;
}

// Slots [count, cap) of a screened-pair buffer get all-ones keys and values,
// so a sort over all cap slots puts the real edges first and unpacks the
// rest as unused (ei = ej = -1).
__global__ void k_pad_tail(unsigned long long* __restrict__ keys,
                           unsigned long long* __restrict__ vals,
                           const unsigned long long* __restrict__ count, long long cap) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    const unsigned long long c = *count;
    if ((unsigned long long)i >= c) {
        keys[i] = ~0ull;
        vals[i] = ~0ull;
    }
}

struct U64Less {
    __device__ __host__ bool operator()(unsigned long long a, unsigned long long b) const {
        return a < b;
    }
};

template <int TILE>
int launch_screen(const float* pts, int64_t n, int d, float thr_sq, uint64_t* keys,
                  uint64_t* vals, int64_t cap, uint64_t* count_dev, hipStream_t s) {
    const long long nb = (n + TILE - 1) / TILE;
    const long long ntiles = nb * (nb + 1) / 2;
    const long long kMaxGrid = 1LL << 30;
    for (long long t0 = 0; t0 < ntiles; t0 += kMaxGrid) {
        const long long g = std::min(kMaxGrid, ntiles - t0);
        k_edge_screen<TILE><<<(unsigned)g, kThreads, 0, s>>>(
            pts, (int)n, d, thr_sq, t0, (unsigned long long*)keys, (unsigned long long*)vals,
            (long long)cap, (unsigned long long*)count_dev);
        FPH_LAUNCH_CHECK();
    }
    return FPH_OK;
}

}  // namespace

extern "C" int fph_edge_screen(const float* pts, int64_t n, int d, float thr_sq,
                               uint64_t* keys, uint64_t* vals, int64_t cap,
                               uint64_t* count_dev, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (n < 0 || d < 1 || cap < 0 || !count_dev) return FPH_ERR_ARG;
    if (n > 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    k_zero_u64<<<1, 64, 0, s>>>((unsigned long long*)count_dev, 1);
    FPH_LAUNCH_CHECK();
    if (n < 2) return FPH_OK;
    switch (fph_tune_value(FPH_TUNE_SCREEN_TILE, n, d)) {
        case 32: return launch_screen<32>(pts, n, d, thr_sq, keys, vals, cap, count_dev, s);
        case 128: return launch_screen<128>(pts, n, d, thr_sq, keys, vals, cap, count_dev, s);
        default: return launch_screen<64>(pts, n, d, thr_sq, keys, vals, cap, count_dev, s);
    }
}

// exact_colex = 0: the reference's order (packed key, ties in emission
// order); 1: (dsq, full colex) -- the two differ only for n > 92682.
static int edge_sort(const uint64_t* keys_in, const uint64_t* vals_in, int64_t E, int64_t n,
                     uint64_t* keys_tmp, uint64_t* vals_tmp, int32_t* edge_i, int32_t* edge_j,
                     float* edge_dsq, float* edge_filt, void* ws, size_t* ws_bytes,
                     hipStream_t s, int exact_colex) {
    if (E < 0 || !ws_bytes) return FPH_ERR_ARG;
    if (E >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    const bool tie_order = n > 92682;
    const unsigned N = (unsigned)E;
    // workspace layout: [sort temp | ord | ord_sorted | idx | idx_sorted]
    size_t tmp_main = 0, tmp_ord = 0;
    FPH_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_main, (const unsigned long long*)keys_in,
                                            (unsigned long long*)keys_tmp,
                                            (const unsigned long long*)vals_in,
                                            (unsigned long long*)vals_tmp, N, 0, 64, s));
    if (tie_order) {
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_ord, (const unsigned long long*)nullptr,
                                                (unsigned long long*)nullptr,
                                                (const unsigned*)nullptr, (unsigned*)nullptr, N,
                                                0, 64, s));
    }
    size_t tmp = align_up(std::max(tmp_main, tmp_ord));
    size_t need = tmp;
    if (tie_order) need += 2 * align_up(8 * (size_t)E) + 2 * align_up(4 * (size_t)E);
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    if (E == 0) return FPH_OK;
    char* w = (char*)ws;
    const unsigned long long* k_src = (const unsigned long long*)keys_in;
    const unsigned long long* v_src = (const unsigned long long*)vals_in;
    unsigned long long* kk = (unsigned long long*)keys_tmp;
    unsigned long long* vv = (unsigned long long*)vals_tmp;
    if (tie_order) {
        unsigned long long* ord = (unsigned long long*)(w + tmp);
        unsigned long long* ord2 = (unsigned long long*)(w + tmp + align_up(8 * (size_t)E));
        unsigned* idx = (unsigned*)(w + tmp + 2 * align_up(8 * (size_t)E));
        unsigned* idx2 = (unsigned*)(w + tmp + 2 * align_up(8 * (size_t)E) + align_up(4 * (size_t)E));
        const long long nb32 = (n + 31) / 32;
        if (exact_colex)
            k_colex_order<<<fph_blocks(E, 256), 256, 0, s>>>(v_src, E, ord, idx);
        else
            k_emit_order<<<fph_blocks(E, 256), 256, 0, s>>>(v_src, E, nb32, ord, idx);
        FPH_LAUNCH_CHECK();
        size_t t2 = tmp;
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, t2, ord, ord2, idx, idx2, N, 0, 64, s));
        // keys/vals in emission order, written to the caller's tmp buffers and
        // sorted back into ord/idx space reinterpreted as u64 pairs below.
        k_gather_kv<<<fph_blocks(E, 256), 256, 0, s>>>(k_src, v_src, idx2, E, kk, vv);
        FPH_LAUNCH_CHECK();
        // Final stable sort on the packed key (emission order breaks ties),
        // or, exact order, on dsq bits << 32 | colex position (unique keys).
        if (exact_colex) {
            k_rank_low<<<fph_blocks(E, 256), 256, 0, s>>>(kk, E);
            FPH_LAUNCH_CHECK();
        }
        unsigned long long* kout = ord;   // reuse (E u64 each)
        unsigned long long* vout = ord2;
        size_t t3 = tmp;
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, t3, kk, kout, vv, vout, N, 0, 64, s));
        k_edge_unpack<<<fph_blocks(E, 256), 256, 0, s>>>(kout, vout, E, edge_i, edge_j, edge_dsq,
                                                         edge_filt);
        FPH_LAUNCH_CHECK();
        return FPH_OK;
    }
    size_t t1 = tmp;
    FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, t1, k_src, kk, v_src, vv, N, 0, 64, s));
    k_edge_unpack<<<fph_blocks(E, 256), 256, 0, s>>>(kk, vv, E, edge_i, edge_j, edge_dsq,
                                                     edge_filt);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_edge_sort(const uint64_t* keys_in, const uint64_t* vals_in, int64_t E,
                             int64_t n, uint64_t* keys_tmp, uint64_t* vals_tmp,
                             int32_t* edge_i, int32_t* edge_j, float* edge_dsq,
                             float* edge_filt, void* ws, size_t* ws_bytes, void* stream) {
    return edge_sort(keys_in, vals_in, E, n, keys_tmp, vals_tmp, edge_i, edge_j, edge_dsq,
                     edge_filt, ws, ws_bytes, (hipStream_t)stream, 0);
}

extern "C" int fph_edge_sort_colex64(const uint64_t* keys_in, const uint64_t* vals_in, int64_t E,
                                     int64_t n, uint64_t* keys_tmp, uint64_t* vals_tmp,
                                     int32_t* edge_i, int32_t* edge_j, float* edge_dsq,
                                     float* edge_filt, void* ws, size_t* ws_bytes, void* stream) {
    return edge_sort(keys_in, vals_in, E, n, keys_tmp, vals_tmp, edge_i, edge_j, edge_dsq,
                     edge_filt, ws, ws_bytes, (hipStream_t)stream, 1);
}

extern "C" int fph_vertex_csr(const int32_t* edge_i, const int32_t* edge_j,
                              const float* edge_dsq, int64_t E, int64_t n, int32_t* adj_ptr,
                              int32_t* adj_idx, int32_t* adj_rank, float* adj_dsq, void* ws,
                              size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || n < 0 || !ws_bytes) return FPH_ERR_ARG;
    if (E >= 0x3FFFFFFFLL) return FPH_ERR_LIMIT;
    const size_t H = 2 * (size_t)E;
    size_t tmp = 0;
    const int end_bit = 32 + std::max(1, bit_width_u64((unsigned long long)std::max<int64_t>(n, 1)));
    FPH_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmp, (const unsigned long long*)nullptr,
                                            (unsigned long long*)nullptr, (const unsigned*)nullptr,
                                            (unsigned*)nullptr, (unsigned)H, 0, end_bit, s));
    tmp = align_up(tmp);
    const size_t need = tmp + 2 * align_up(8 * H) + 2 * align_up(4 * H);
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    unsigned long long* hk = (unsigned long long*)(w + tmp);
    unsigned long long* hk2 = (unsigned long long*)(w + tmp + align_up(8 * H));
    unsigned* hv = (unsigned*)(w + tmp + 2 * align_up(8 * H));
    unsigned* hv2 = (unsigned*)(w + tmp + 2 * align_up(8 * H) + align_up(4 * H));
    if (E > 0) {
        k_half_edges<<<fph_blocks(E, 256), 256, 0, s>>>(edge_i, edge_j, E, hk, hv);
        FPH_LAUNCH_CHECK();
        size_t t1 = tmp;
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, t1, hk, hk2, hv, hv2, (unsigned)H, 0, end_bit, s));
        k_csr_unpack<<<fph_blocks((int64_t)H, 256), 256, 0, s>>>(hk2, hv2, (long long)H, edge_dsq,
                                                                adj_idx, adj_rank, adj_dsq);
        FPH_LAUNCH_CHECK();
    }
    k_csr_ptr<<<fph_blocks(n + 1, 256), 256, 0, s>>>(hk2, (long long)H, n, adj_ptr);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" const char* fph_version(void) { return "flashph-mi355x 0.1.0 (gfx950)"; }

// ---------------------------------------------------------------------------
// Capacity form for graph capture: every size is the capacity `cap` fixed on
// the host, the live edge count is read on the device, and nothing records a
// memset node or synchronises (sorts use rocprim's merge sort, whose launches
// are kernels only).  Requires n <= 92682, where the packed keys are unique
// and the sort order needs no emission-order tie-break.
// ---------------------------------------------------------------------------

int fph_edges_csr_capacity(uint64_t* keys, uint64_t* vals, const uint64_t* count_dev,
                           int64_t cap, int64_t n, int32_t* edge_i, int32_t* edge_j,
                           float* edge_dsq, float* edge_filt, int32_t* adj_ptr, int32_t* adj_idx,
                           int32_t* adj_rank, float* adj_dsq, void* ws, size_t* ws_bytes,
                           hipStream_t s) {
    if (cap < 1 || n < 1 || !ws_bytes) return FPH_ERR_ARG;
    if (n > 92682 || cap >= 0x3FFFFFFFLL) return FPH_ERR_LIMIT;
    const size_t H = 2 * (size_t)cap;
    size_t t_e = 0, t_h = 0;
    using K = unsigned long long;
    FPH_CHECK_HIP(rocprim::merge_sort(nullptr, t_e, (K*)nullptr, (K*)nullptr, (K*)nullptr,
                                      (K*)nullptr, (size_t)cap, U64Less(), s));
    FPH_CHECK_HIP(rocprim::merge_sort(nullptr, t_h, (K*)nullptr, (K*)nullptr, (unsigned*)nullptr,
                                      (unsigned*)nullptr, H, U64Less(), s));
    const size_t tmp = align_up(std::max(t_e, t_h));
    // [sort temp | sorted keys | sorted vals | hk | hk2 | hv | hv2]
    const size_t need = tmp + 2 * align_up(8 * (size_t)cap) + 2 * align_up(8 * H) +
                        2 * align_up(4 * H);
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    K* ks = (K*)(w + tmp);
    K* vs = (K*)(w + tmp + align_up(8 * (size_t)cap));
    char* w2 = w + tmp + 2 * align_up(8 * (size_t)cap);
    K* hk = (K*)w2;
    K* hk2 = (K*)(w2 + align_up(8 * H));
    unsigned* hv = (unsigned*)(w2 + 2 * align_up(8 * H));
    unsigned* hv2 = (unsigned*)(w2 + 2 * align_up(8 * H) + align_up(4 * H));
    k_pad_tail<<<fph_blocks(cap, 256), 256, 0, s>>>((K*)keys, (K*)vals, (const K*)count_dev, cap);
    FPH_LAUNCH_CHECK();
    size_t t1 = tmp;
    FPH_CHECK_HIP(rocprim::merge_sort(ws, t1, (K*)keys, ks, (K*)vals, vs, (size_t)cap, U64Less(), s));
    k_edge_unpack<<<fph_blocks(cap, 256), 256, 0, s>>>(ks, vs, cap, edge_i, edge_j, edge_dsq,
                                                       edge_filt);
    FPH_LAUNCH_CHECK();
    k_half_edges<<<fph_blocks(cap, 256), 256, 0, s>>>(edge_i, edge_j, cap, hk, hv);
    FPH_LAUNCH_CHECK();
    size_t t2 = tmp;
    FPH_CHECK_HIP(rocprim::merge_sort(ws, t2, hk, hk2, hv, hv2, H, U64Less(), s));
    k_csr_unpack<<<fph_blocks((int64_t)H, 256), 256, 0, s>>>(hk2, hv2, (long long)H, edge_dsq,
                                                            adj_idx, adj_rank, adj_dsq);
    FPH_LAUNCH_CHECK();
    k_csr_ptr<<<fph_blocks(n + 1, 256), 256, 0, s>>>(hk2, (long long)H, n, adj_ptr);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
