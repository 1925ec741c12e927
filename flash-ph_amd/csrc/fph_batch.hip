// Batches of point clouds of uneven size and dimension in one pass.
//
// A batch of disjoint clouds is one Vietoris-Rips complex whose pairing is the
// union of the clouds' pairings, as long as each cloud keeps its own edge
// order: edges are screened per cloud (own n, d, threshold, local colex key),
// sorted by (cloud, key), and vertices get global ids (cloud base + local
// id).  From the CSR on, the single-cloud stages run unchanged on the union;
// the reduction and the diagrams split the columns back into clouds.
//
// Also here: the k-nearest-neighbour threshold of the reference's
// auto_threshold (threshold.py:20-78, 133-176), per cloud.
#include "fph_common.h"

#include <rocprim/rocprim.hpp>

namespace {

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

constexpr int kTile = 64;
constexpr int kChunk = 32;
constexpr int kThreads = 256;

__device__ __forceinline__ int find_seg(const long long* off, int nseg, long long x) {
    int lo = 0, hi = nseg;  // last s with off[s] <= x
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void tile_coords(long long t, int* I, int* J) {
    long long j = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= t) j++;
    while (j * (j + 1) / 2 > t) j--;
    *J = (int)j;
    *I = (int)(t - j * (j + 1) / 2);
}

__global__ void k_zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0) p[0] = 0ull;
}

// Same tile computation as the single-cloud screen (64 x 64 tiles, FMA chain
// in dimension order, strict FP32 threshold), over the concatenated tile
// list of all clouds.  key = float_bits(dsq) << 32 | local colex mod 2^32,
// val = global j << 32 | global i.
__global__ __launch_bounds__(kThreads) void k_edge_screen_batch(
    const float* __restrict__ pts, const long long* __restrict__ foff,
    const int* __restrict__ dims, const long long* __restrict__ vbase,
    const float* __restrict__ thr_sq, const long long* __restrict__ tile_off, int C,
    long long tile0, unsigned long long* __restrict__ keys, unsigned long long* __restrict__ vals,
    long long cap, unsigned long long* __restrict__ count) {
    __shared__ float4 sa[kChunk][kTile / 4];
    __shared__ float4 sb[kChunk][kTile / 4];
    const long long t = tile0 + blockIdx.x;
    const int c = find_seg(tile_off, C, t);
    int I, J;
    tile_coords(t - tile_off[c], &I, &J);
    const int n = (int)(vbase[c + 1] - vbase[c]);
    const int d = dims[c];
    const float thr = thr_sq[c];
    const float* P = pts + foff[c];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = I * kTile, j0 = J * kTile;
    float acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[r][q] = 0.0f;
    float* fa = reinterpret_cast<float*>(sa);
    float* fb = reinterpret_cast<float*>(sb);
    for (int k0 = 0; k0 < d; k0 += kChunk) {
        const int kc = min(kChunk, d - k0);
        __syncthreads();
        for (int idx = threadIdx.x; idx < kTile * kc; idx += kThreads) {
            const int p = idx / kc, k = idx - p * kc;
            const int gi = i0 + p, gj = j0 + p;
            fa[k * kTile + p] = gi < n ? P[(long long)gi * d + k0 + k] : 0.0f;
            fb[k * kTile + p] = gj < n ? P[(long long)gj * d + k0 + k] : 0.0f;
        }
        __syncthreads();
        for (int k = 0; k < kc; k++) {
            const float4 a = sa[k][ty];
            const float4 b = sb[k][tx];
            const float av[4] = {a.x, a.y, a.z, a.w};
            const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float diff = av[r] - bv[q];
                    acc[r][q] = __builtin_fmaf(diff, diff, acc[r][q]);
                }
        }
    }
    unsigned long long bal[16];
    int total = 0;
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = i0 + ty * 4 + r, j = j0 + tx * 4 + q;
            const bool ok = (acc[r][q] < thr) && i < n && j < n && (I < J || i < j);
            bal[r * 4 + q] = __ballot(ok);
            total += __popcll(bal[r * 4 + q]);
        }
    unsigned long long pos = fph_wg_reserve<kThreads>(total, count);
    if (total == 0) return;
    const unsigned long long lt = fph_lanemask_lt();
    const unsigned long long vb = (unsigned long long)vbase[c];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const unsigned long long b = bal[r * 4 + q];
            if ((b >> fph_lane()) & 1ull) {
                const unsigned long long at = pos + __popcll(b & lt);
                if (at < (unsigned long long)cap) {
                    const unsigned long long i = (unsigned long long)(i0 + ty * 4 + r);
                    const unsigned long long j = (unsigned long long)(j0 + tx * 4 + q);
                    unsigned bits;
                    memcpy(&bits, &acc[r][q], 4);
                    const unsigned long long colex = j * (j - 1) / 2 + i;
                    keys[at] = ((unsigned long long)bits << 32) | (colex & 0xFFFFFFFFull);
                    vals[at] = ((vb + j) << 32) | (vb + i);
                }
            }
            pos += __popcll(b);
        }
}

__global__ void k_cloud_of_edge(const unsigned long long* __restrict__ vals, long long E,
                                const long long* __restrict__ vbase, int C,
                                unsigned* __restrict__ cid, unsigned* __restrict__ idx) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    cid[e] = (unsigned)find_seg(vbase, C, (long long)(vals[e] & 0xFFFFFFFFull));
    idx[e] = (unsigned)e;
}

__global__ void k_gather_unpack(const unsigned long long* __restrict__ keys,
                                const unsigned long long* __restrict__ vals,
                                const unsigned* __restrict__ perm, long long E,
                                int* __restrict__ ei, int* __restrict__ ej,
                                float* __restrict__ dsq, float* __restrict__ filt) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const unsigned s = perm[e];
    const unsigned long long v = vals[s];
    const unsigned bits = (unsigned)(keys[s] >> 32);
    float f;
    memcpy(&f, &bits, 4);
    ei[e] = (int)(v & 0xFFFFFFFFull);
    ej[e] = (int)(v >> 32);
    dsq[e] = f;
    filt[e] = fph_sqrt_rn(f);
}

__global__ void k_edge_offsets(const unsigned* __restrict__ cid_sorted, long long E, int C,
                               int* __restrict__ edge_off) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > C) return;
    long long lo = 0, hi = E;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if ((int)cid_sorted[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    edge_off[c] = (int)lo;
}

// ---------------------------------------------------------------- H0

__global__ void k_mask_i32(const unsigned char* __restrict__ m, long long E, int* __restrict__ o) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) o[e] = m[e];
}

// pos[e] = forest edges before e (exclusive scan); pos[E] = total.
__global__ void k_h0_batch(const unsigned char* __restrict__ mst, const int* __restrict__ pos,
                           const float* __restrict__ dsq, long long E,
                           const int* __restrict__ edge_off, const long long* __restrict__ vbase,
                           int C, long long N, float* __restrict__ h0) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < E && mst[i]) {
        int lo = 0, hi = C;  // cloud of edge i: last c with edge_off[c] <= i
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (edge_off[mid] <= i) lo = mid;
            else hi = mid;
        }
        const long long row = vbase[lo] + pos[i] - pos[edge_off[lo]];
        if (row < vbase[lo + 1]) {
            h0[2 * row] = 0.0f;
            h0[2 * row + 1] = fph_sqrt_rn(dsq[i]);
        }
    }
    if (i < N) {
        const int c = find_seg(vbase, C, i);
        const long long nm = pos[edge_off[c + 1]] - pos[edge_off[c]];
        if (i - vbase[c] >= nm) {
            h0[2 * i] = 0.0f;
            h0[2 * i + 1] = __builtin_inff();
        }
    }
}

__global__ void k_scan_total(const int* __restrict__ m, int* __restrict__ pos, long long E) {
    if (threadIdx.x == 0 && blockIdx.x == 0) pos[E] = E > 0 ? pos[E - 1] + m[E - 1] : 0;
}

// ---------------------------------------------------------------- kNN

constexpr int kKnnThreads = 256;
constexpr int kKnnCand = 64;      // candidates staged per LDS pass (bucketed dims)
constexpr int kKnnLdsFloats = 12288;  // 48 KB: staging for the generic (d > 64) path

// Insert v into a descending list holding the k smallest values seen so far
// (list[0] is the current k-th smallest; slots >= k hold -inf and never
// change): drops list[0] and places v in order.
template <int KM>
__device__ __forceinline__ void knn_insert(float (&top)[KM], float v) {
#pragma unroll
    for (int x = 0; x < KM - 1; x++) top[x] = fmaxf(top[x + 1], fminf(top[x], v));
    top[KM - 1] = fminf(top[KM - 1], v);
}

// Dimensions padded with zeros up to DP: the padded terms add an exact 0 at
// the end of the FMA chain, so the sum equals the screen's.
template <int DP, int KM>
__device__ __forceinline__ float knn_scan(const float* __restrict__ P, int n, int d, int q,
                                          bool active, int k, float* cand) {
    float xq[DP];
#pragma unroll
    for (int x = 0; x < DP; x++) xq[x] = (active && x < d) ? P[(long long)q * d + x] : 0.0f;
    float top[KM];
#pragma unroll
    for (int x = 0; x < KM; x++) top[x] = x < k ? __builtin_inff() : -__builtin_inff();
    for (int j0 = 0; j0 < n; j0 += kKnnCand) {
        const int nc = min(kKnnCand, n - j0);
        __syncthreads();
        for (int idx = threadIdx.x; idx < nc * DP; idx += kKnnThreads) {
            const int jj = idx / DP, x = idx % DP;
            cand[idx] = x < d ? P[(long long)(j0 + jj) * d + x] : 0.0f;
        }
        __syncthreads();
        for (int jj = 0; jj < nc; jj++) {
            const float* y = cand + jj * DP;
            float acc = 0.0f;
#pragma unroll
            for (int x = 0; x < DP; x++) {
                const float diff = xq[x] - y[x];
                acc = __builtin_fmaf(diff, diff, acc);
            }
            acc = (j0 + jj == q) ? __builtin_inff() : acc;
            if (acc < top[0]) knn_insert<KM>(top, acc);
        }
    }
    return top[0];
}

template <int KM>
__device__ __forceinline__ float knn_scan_generic(const float* __restrict__ P, int n, int d,
                                                  int q, bool active, int k, float* cand) {
    float top[KM];
#pragma unroll
    for (int x = 0; x < KM; x++) top[x] = x < k ? __builtin_inff() : -__builtin_inff();
    const int per = max(1, kKnnLdsFloats / d);
    const float* xq = P + (long long)(active ? q : 0) * d;
    for (int j0 = 0; j0 < n; j0 += per) {
        const int nc = min(per, n - j0);
        __syncthreads();
        for (int idx = threadIdx.x; idx < nc * d; idx += kKnnThreads)
            cand[idx] = P[(long long)j0 * d + idx];
        __syncthreads();
        for (int jj = 0; jj < nc; jj++) {
            const float* y = cand + jj * d;
            float acc = 0.0f;
            for (int x = 0; x < d; x++) {
                const float diff = xq[x] - y[x];
                acc = __builtin_fmaf(diff, diff, acc);
            }
            acc = (j0 + jj == q) ? __builtin_inff() : acc;
            if (acc < top[0]) knn_insert<KM>(top, acc);
        }
    }
    return top[0];
}

// One thread per query point, 256 queries of one cloud per block: squared
// distance (the screen's FMA chain) to every other point of the cloud, the
// k smallest kept in registers; writes the k-th smallest squared distance.
// Matches the k-th entry of the reference's per-row top-k over a distance
// matrix with the diagonal masked (threshold.py:133-176), with the distance
// taken in the filtration's own FP32 form.
template <int KM>
__global__ __launch_bounds__(kKnnThreads) void k_knn_batch(
    const float* __restrict__ pts, const long long* __restrict__ foff,
    const int* __restrict__ dims, const long long* __restrict__ vbase, int C, int k_req,
    const long long* __restrict__ qtile_off, float* __restrict__ kth) {
    extern __shared__ float cand[];
    const long long t = blockIdx.x;
    const int c = find_seg(qtile_off, C, t);
    const int n = (int)(vbase[c + 1] - vbase[c]);
    const int d = dims[c];
    const float* P = pts + foff[c];
    const int q = (int)(t - qtile_off[c]) * kKnnThreads + threadIdx.x;
    const bool active = q < n;
    const int k = max(1, min(k_req, n - 1));
    float r;
    if (d <= 4) r = knn_scan<4, KM>(P, n, d, q, active, k, cand);
    else if (d <= 8) r = knn_scan<8, KM>(P, n, d, q, active, k, cand);
    else if (d <= 16) r = knn_scan<16, KM>(P, n, d, q, active, k, cand);
    else if (d <= 32) r = knn_scan<32, KM>(P, n, d, q, active, k, cand);
    else if (d <= 64) r = knn_scan<64, KM>(P, n, d, q, active, k, cand);
    else r = knn_scan_generic<KM>(P, n, d, q, active, k, cand);
    if (active) kth[vbase[c] + q] = n > 1 ? r : 0.0f;
}

// k-th smallest squared distance for k beyond the register lists
// (k > FPH_KNN_MAX_K): one wavefront per query point selects it exactly by
// radix on the FP32 bits (non-negative floats order as their bit patterns):
// four passes over the cloud, each histogramming the next 8 bits of the
// distances that carry the bits fixed so far (256 LDS counters per
// wavefront) and fixing the byte where the cumulative count reaches the
// remaining rank.  Distances are the same FMA chain as k_knn_batch, so the
// result is the k-th entry of the same sorted list.
constexpr int kSelWaves = 16;

__global__ __launch_bounds__(kSelWaves * FPH_WAVE) void k_knn_select(
    const float* __restrict__ pts, const long long* __restrict__ foff,
    const int* __restrict__ dims, const long long* __restrict__ vbase, int C, long long N,
    int k_req, float* __restrict__ kth) {
    __shared__ unsigned hist[kSelWaves][256];
    const int wv = threadIdx.x / FPH_WAVE, lane = fph_lane();
    const long long g = (long long)blockIdx.x * kSelWaves + wv;  // global query id
    if (g >= N) return;                                            // wave-uniform
    const int c = find_seg(vbase, C, g);
    const int n = (int)(vbase[c + 1] - vbase[c]);
    const int q = (int)(g - vbase[c]);
    const int d = dims[c];
    const float* P = pts + foff[c];
    const float* xq = P + (long long)q * d;
    if (n <= 1) {
        if (lane == 0) kth[g] = 0.0f;
        return;
    }
    unsigned krem = (unsigned)max(1, min(k_req, n - 1));  // rank among the n-1 others
    unsigned prefix = 0;
    unsigned* h = hist[wv];
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        for (int b = lane; b < 256; b += FPH_WAVE) h[b] = 0u;
        __builtin_amdgcn_wave_barrier();
        for (int j = lane; j < n; j += FPH_WAVE) {
            if (j == q) continue;
            const float* y = P + (long long)j * d;
            float acc = 0.0f;
            for (int x = 0; x < d; x++) {
                const float diff = xq[x] - y[x];
                acc = __builtin_fmaf(diff, diff, acc);
            }
            unsigned bits;
            memcpy(&bits, &acc, 4);
            if (pass == 0 || (bits >> (shift + 8)) == prefix)
                atomicAdd(&h[(bits >> shift) & 255u], 1u);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        // lane l holds bins 4l .. 4l+3; inclusive scan of their sums
        unsigned c4[4], sum = 0;
        for (int t = 0; t < 4; t++) {
            c4[t] = h[4 * lane + t];
            sum += c4[t];
        }
        unsigned incl = sum;
        for (int o = 1; o < FPH_WAVE; o <<= 1) {
            const unsigned y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const unsigned excl = incl - sum;
        // the lane whose range reaches rank krem picks the byte
        const bool mine = excl < krem && incl >= krem;
        const unsigned long long who = __ballot(mine);
        const int src = who ? __ffsll((long long)who) - 1 : 0;  // always found
        unsigned byte = 0, before = excl;
        if (mine) {
            for (int t = 0; t < 4; t++) {
                if (before + c4[t] >= krem) { byte = 4 * lane + t; break; }
                before += c4[t];
            }
        }
        byte = (unsigned)__shfl((int)byte, src);
        before = (unsigned)__shfl((int)before, src);
        prefix = (prefix << 8) | byte;
        krem -= before;
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) {
        float v;
        memcpy(&v, &prefix, 4);
        kth[g] = v;
    }
}

// Largest squared distance from each point to the rest of its cloud
// (threshold.py:179-200 takes the minimum of these over the cloud); the
// cloud's minimum goes to rmin[c] as FP32 bits (non-negative floats order
// as their bit patterns, so an unsigned atomicMin is exact).
template <int DP>
__device__ __forceinline__ float far_scan(const float* __restrict__ P, int n, int d, int q,
                                          bool active, float* cand) {
    float xq[DP];
#pragma unroll
    for (int x = 0; x < DP; x++) xq[x] = (active && x < d) ? P[(long long)q * d + x] : 0.0f;
    float best = 0.0f;
    for (int j0 = 0; j0 < n; j0 += kKnnCand) {
        const int nc = min(kKnnCand, n - j0);
        __syncthreads();
        for (int idx = threadIdx.x; idx < nc * DP; idx += kKnnThreads) {
            const int jj = idx / DP, x = idx % DP;
            cand[idx] = x < d ? P[(long long)(j0 + jj) * d + x] : 0.0f;
        }
        __syncthreads();
        for (int jj = 0; jj < nc; jj++) {
            const float* y = cand + jj * DP;
            float acc = 0.0f;
#pragma unroll
            for (int x = 0; x < DP; x++) {
                const float diff = xq[x] - y[x];
                acc = __builtin_fmaf(diff, diff, acc);
            }
            best = fmaxf(best, acc);
        }
    }
    return best;
}

__device__ __forceinline__ float far_scan_generic(const float* __restrict__ P, int n, int d,
                                                  int q, bool active, float* cand) {
    float best = 0.0f;
    const int per = max(1, kKnnLdsFloats / d);
    const float* xq = P + (long long)(active ? q : 0) * d;
    for (int j0 = 0; j0 < n; j0 += per) {
        const int nc = min(per, n - j0);
        __syncthreads();
        for (int idx = threadIdx.x; idx < nc * d; idx += kKnnThreads)
            cand[idx] = P[(long long)j0 * d + idx];
        __syncthreads();
        for (int jj = 0; jj < nc; jj++) {
            const float* y = cand + jj * d;
            float acc = 0.0f;
            for (int x = 0; x < d; x++) {
                const float diff = xq[x] - y[x];
                acc = __builtin_fmaf(diff, diff, acc);
            }
            best = fmaxf(best, acc);
        }
    }
    return best;
}

__global__ __launch_bounds__(kKnnThreads) void k_enclosing_batch(
    const float* __restrict__ pts, const long long* __restrict__ foff,
    const int* __restrict__ dims, const long long* __restrict__ vbase, int C,
    const long long* __restrict__ qtile_off, unsigned* __restrict__ rmin) {
    extern __shared__ float cand[];
    const long long t = blockIdx.x;
    const int c = find_seg(qtile_off, C, t);
    const int n = (int)(vbase[c + 1] - vbase[c]);
    const int d = dims[c];
    const float* P = pts + foff[c];
    const int q = (int)(t - qtile_off[c]) * kKnnThreads + threadIdx.x;
    const bool active = q < n;
    float r;
    if (d <= 4) r = far_scan<4>(P, n, d, q, active, cand);
    else if (d <= 8) r = far_scan<8>(P, n, d, q, active, cand);
    else if (d <= 16) r = far_scan<16>(P, n, d, q, active, cand);
    else if (d <= 32) r = far_scan<32>(P, n, d, q, active, cand);
    else if (d <= 64) r = far_scan<64>(P, n, d, q, active, cand);
    else r = far_scan_generic(P, n, d, q, active, cand);
    unsigned b;
    memcpy(&b, &r, 4);
    b = active ? b : 0xFFFFFFFFu;
    // wave minimum, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) b = min(b, (unsigned)__shfl_xor((int)b, o));
    if (fph_lane() == 0 && b != 0xFFFFFFFFu) atomicMin(&rmin[c], b);
}

__global__ void k_fill_u32(unsigned* p, long long n, unsigned v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// r_out[c] = sqrt of the cloud's minimum (0 for n <= 1, as threshold.py:107).
__global__ void k_enclosing_finish(const unsigned* __restrict__ rmin,
                                   const long long* __restrict__ vbase, int C,
                                   float* __restrict__ r_out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float v;
    const unsigned b = rmin[c];
    memcpy(&v, &b, 4);
    r_out[c] = (vbase[c + 1] - vbase[c] > 1) ? fph_sqrt_rn(v) : 0.0f;
}

__global__ void k_sqrt_bits(const float* __restrict__ dsq, long long N, unsigned* __restrict__ o,
                            float* __restrict__ dist) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float v = fph_sqrt_rn(dsq[i]);
    unsigned b;
    memcpy(&b, &v, 4);
    o[i] = b;  // non-negative floats sort as their bits
    if (dist) dist[i] = v;
}

// torch.quantile(x, q) with linear interpolation, in FP32 as torch does for
// float input: rank = q * (n - 1), lerp between the neighbouring order
// statistics (torch's lerp: a + w (b - a) for w < 0.5, b - (b - a)(1 - w)).
// Clouds of n <= 1 give 0 (threshold.py:59-61).  thr_sq = fp32(t * t) with
// t * t in double (the Python float product of rips.py:125).
__global__ void k_quantile(const unsigned* __restrict__ sorted_bits,
                           const long long* __restrict__ vbase, int C, float qf,
                           float* __restrict__ t_out, float* __restrict__ thr_sq_out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const long long n = vbase[c + 1] - vbase[c];
    float t = 0.0f;
    if (n > 1) {
        const float rank = qf * (float)(n - 1);
        const float lo_r = floorf(rank), hi_r = ceilf(rank);
        const float w = rank - lo_r;
        float a, b;
        const unsigned ab = sorted_bits[vbase[c] + (long long)lo_r];
        const unsigned bb = sorted_bits[vbase[c] + (long long)hi_r];
        memcpy(&a, &ab, 4);
        memcpy(&b, &bb, 4);
        t = w < 0.5f ? a + w * (b - a) : b - (b - a) * (1.0f - w);
    }
    t_out[c] = t;
    if (thr_sq_out) thr_sq_out[c] = (float)((double)t * (double)t);
}

}  // namespace

// ------------------------------------------------------------------- C-ABI

template <int KM>
static hipError_t launch_knn(const float* pts, const int64_t* foff, const int32_t* dims,
                             const int64_t* vbase, int C, int k, const int64_t* qtile_off,
                             int64_t n_qtiles, size_t lds, float* kth, hipStream_t s) {
    k_knn_batch<KM><<<(unsigned)n_qtiles, kKnnThreads, lds, s>>>(
        pts, (const long long*)foff, dims, (const long long*)vbase, C, k,
        (const long long*)qtile_off, kth);
    return hipGetLastError();
}

extern "C" int fph_knn_threshold_batch(const float* pts, const int64_t* foff, const int32_t* dims,
                                       const int64_t* vbase, const int64_t* qtile_off,
                                       int64_t n_qtiles, int64_t C, int64_t N, int k,
                                       double percentile, int max_dim_host, float* kth_dist,
                                       float* t_out, float* thr_sq_out, void* ws,
                                       size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (C < 1 || N < 0 || k < 1 || !ws_bytes || max_dim_host < 1 ||
        n_qtiles < 0 || n_qtiles > 0x7FFFFFFFLL || N >= 0x7FFFFFFFLL)
        return FPH_ERR_ARG;
    const size_t n4 = align_up(4 * (size_t)std::max<int64_t>(N, 1));
    size_t tmp = 0;
    FPH_CHECK_HIP(rocprim::segmented_radix_sort_keys(
        nullptr, tmp, (const unsigned*)nullptr, (unsigned*)nullptr,
        (unsigned)std::max<int64_t>(N, 1), (unsigned)C, (const long long*)nullptr,
        (const long long*)nullptr, 0, 32, s));
    tmp = align_up(tmp);
    const size_t need = tmp + 3 * n4;
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    float* kth = (float*)(w + tmp);
    unsigned* bits = (unsigned*)(w + tmp + n4);
    unsigned* sorted = (unsigned*)(w + tmp + 2 * n4);
    if (N > 0 && n_qtiles > 0) {
        const size_t lds = 4 * (size_t)(max_dim_host > 64 ? kKnnLdsFloats : kKnnCand * 64);
        hipError_t e;
        if (k > FPH_KNN_MAX_K) {
            k_knn_select<<<(unsigned)((N + kSelWaves - 1) / kSelWaves), kSelWaves * FPH_WAVE, 0,
                           s>>>(pts, (const long long*)foff, dims, (const long long*)vbase,
                                (int)C, N, k, kth);
            e = hipGetLastError();
        } else if (k <= 16) e = launch_knn<16>(pts, foff, dims, vbase, (int)C, k, qtile_off, n_qtiles, lds, kth, s);
        else if (k <= 32) e = launch_knn<32>(pts, foff, dims, vbase, (int)C, k, qtile_off, n_qtiles, lds, kth, s);
        else if (k <= 64) e = launch_knn<64>(pts, foff, dims, vbase, (int)C, k, qtile_off, n_qtiles, lds, kth, s);
        else e = launch_knn<128>(pts, foff, dims, vbase, (int)C, k, qtile_off, n_qtiles, lds, kth, s);
        if (e != hipSuccess) return FPH_ERR_HIP;
        k_sqrt_bits<<<fph_blocks(N, 256), 256, 0, s>>>(kth, N, bits, kth_dist);
        FPH_LAUNCH_CHECK();
        size_t t = tmp;
        FPH_CHECK_HIP(rocprim::segmented_radix_sort_keys(ws, t, bits, sorted, (unsigned)N,
                                                         (unsigned)C, (const long long*)vbase,
                                                         (const long long*)vbase + 1, 0, 32, s));
    }
    const float qf = (float)(std::max(0.0, std::min(100.0, percentile)) / 100.0);
    k_quantile<<<fph_blocks(C, 256), 256, 0, s>>>(sorted, (const long long*)vbase, (int)C, qf,
                                                 t_out, thr_sq_out);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_enclosing_radius_batch(const float* pts, const int64_t* foff,
                                          const int32_t* dims, const int64_t* vbase,
                                          const int64_t* qtile_off, int64_t n_qtiles, int64_t C,
                                          int max_dim_host, float* r_out, void* ws,
                                          size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (C < 1 || max_dim_host < 1 || !ws_bytes || n_qtiles < 0 || n_qtiles > 0x7FFFFFFFLL)
        return FPH_ERR_ARG;
    const size_t need = align_up(4 * (size_t)C);
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    unsigned* rmin = (unsigned*)ws;
    k_fill_u32<<<fph_blocks(C, 256), 256, 0, s>>>(rmin, C, 0x7F800000u);  // +inf
    FPH_LAUNCH_CHECK();
    if (n_qtiles > 0) {
        const size_t lds = 4 * (size_t)(max_dim_host > 64 ? kKnnLdsFloats : kKnnCand * 64);
        k_enclosing_batch<<<(unsigned)n_qtiles, kKnnThreads, lds, s>>>(
            pts, (const long long*)foff, dims, (const long long*)vbase, (int)C,
            (const long long*)qtile_off, rmin);
        FPH_LAUNCH_CHECK();
    }
    k_enclosing_finish<<<fph_blocks(C, 256), 256, 0, s>>>(rmin, (const long long*)vbase, (int)C,
                                                         r_out);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_edge_screen_batch(const float* pts, const int64_t* foff, const int32_t* dims,
                                     const int64_t* vbase, const float* thr_sq,
                                     const int64_t* tile_off, int64_t n_tiles, int64_t C,
                                     uint64_t* keys, uint64_t* vals, int64_t cap,
                                     uint64_t* count_dev, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (C < 1 || n_tiles < 0 || cap < 0 || !count_dev) return FPH_ERR_ARG;
    k_zero_u64<<<1, 64, 0, s>>>((unsigned long long*)count_dev);
    FPH_LAUNCH_CHECK();
    const long long kMaxGrid = 1LL << 30;
    for (long long t0 = 0; t0 < n_tiles; t0 += kMaxGrid) {
        const long long g = std::min(kMaxGrid, n_tiles - t0);
        k_edge_screen_batch<<<(unsigned)g, kThreads, 0, s>>>(
            pts, (const long long*)foff, dims, (const long long*)vbase, thr_sq,
            (const long long*)tile_off, (int)C, t0, (unsigned long long*)keys,
            (unsigned long long*)vals, cap, (unsigned long long*)count_dev);
        FPH_LAUNCH_CHECK();
    }
    return FPH_OK;
}

extern "C" int fph_edge_sort_batch(const uint64_t* keys_in, const uint64_t* vals_in, int64_t E,
                                   const int64_t* vbase, int64_t C, int32_t* edge_i,
                                   int32_t* edge_j, float* edge_dsq, float* edge_filt,
                                   int32_t* edge_off, void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || C < 1 || !ws_bytes) return FPH_ERR_ARG;
    if (E >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    const size_t e8 = align_up(8 * (size_t)std::max<int64_t>(E, 1));
    const size_t e4 = align_up(4 * (size_t)std::max<int64_t>(E, 1));
    int cbits = 1;
    while ((1LL << cbits) < C) cbits++;
    size_t t1 = 0, t2 = 0;
    FPH_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, t1, (const unsigned long long*)nullptr,
                                            (unsigned long long*)nullptr,
                                            (const unsigned long long*)nullptr,
                                            (unsigned long long*)nullptr,
                                            (unsigned)std::max<int64_t>(E, 1), 0, 64, s));
    FPH_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, t2, (const unsigned*)nullptr,
                                            (unsigned*)nullptr, (const unsigned*)nullptr,
                                            (unsigned*)nullptr, (unsigned)std::max<int64_t>(E, 1),
                                            0, cbits, s));
    const size_t tmp = align_up(std::max(t1, t2));
    const size_t need = tmp + 2 * e8 + 4 * e4;
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    unsigned long long* ks = (unsigned long long*)(w + tmp);
    unsigned long long* vs = (unsigned long long*)(w + tmp + e8);
    unsigned* cid = (unsigned*)(w + tmp + 2 * e8);
    unsigned* idx = (unsigned*)(w + tmp + 2 * e8 + e4);
    unsigned* cid2 = (unsigned*)(w + tmp + 2 * e8 + 2 * e4);
    unsigned* perm = (unsigned*)(w + tmp + 2 * e8 + 3 * e4);
    if (E > 0) {
        size_t a = tmp;
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, a, (const unsigned long long*)keys_in, ks,
                                                (const unsigned long long*)vals_in, vs,
                                                (unsigned)E, 0, 64, s));
        k_cloud_of_edge<<<fph_blocks(E, 256), 256, 0, s>>>(vs, E, (const long long*)vbase,
                                                          (int)C, cid, idx);
        FPH_LAUNCH_CHECK();
        size_t b = tmp;  // stable: key order is kept inside each cloud
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, b, cid, cid2, idx, perm, (unsigned)E, 0, cbits,
                                                s));
        k_gather_unpack<<<fph_blocks(E, 256), 256, 0, s>>>(ks, vs, perm, E, edge_i, edge_j,
                                                          edge_dsq, edge_filt);
        FPH_LAUNCH_CHECK();
    }
    k_edge_offsets<<<fph_blocks(C + 1, 256), 256, 0, s>>>(cid2, E, (int)C, edge_off);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_h0_diagram_batch(const uint8_t* mst_mask, const float* edge_dsq, int64_t E,
                                    const int32_t* edge_off, const int64_t* vbase, int64_t C,
                                    int64_t N, float* h0, void* ws, size_t* ws_bytes,
                                    void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || C < 1 || N < 0 || !ws_bytes) return FPH_ERR_ARG;
    const size_t e4 = align_up(4 * (size_t)(std::max<int64_t>(E, 1) + 1));
    size_t tmp = 0;
    FPH_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp, (const int*)nullptr, (int*)nullptr, 0,
                                          (size_t)std::max<int64_t>(E, 1), rocprim::plus<int>(), s));
    tmp = align_up(tmp);
    const size_t need = tmp + 2 * e4;
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    int* m32 = (int*)(w + tmp);
    int* pos = (int*)(w + tmp + e4);
    if (E > 0) {
        k_mask_i32<<<fph_blocks(E, 256), 256, 0, s>>>(mst_mask, E, m32);
        FPH_LAUNCH_CHECK();
        size_t t = tmp;
        FPH_CHECK_HIP(rocprim::exclusive_scan(ws, t, m32, pos, 0, (size_t)E, rocprim::plus<int>(), s));
    }
    k_scan_total<<<1, 64, 0, s>>>(m32, pos, E);
    FPH_LAUNCH_CHECK();
    const long long span = std::max<long long>(E, N);
    k_h0_batch<<<fph_blocks(span, 256), 256, 0, s>>>(mst_mask, pos, edge_dsq, E, edge_off,
                                                    (const long long*)vbase, (int)C, N, h0);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
