// Stage (a) on the matrix cores, for wide points (d >= 32).
//
// |x - y|^2 = |x|^2 + |y|^2 - 2 x.y, and the Gram products x.y of 128 x 128
// point tiles run on v_mfma_f32_32x32x16_bf16 with every coordinate split
// into two bf16 parts, x = xh + xl (+ r):
//     x.y ~ xh.yh + xh.yl + xl.yh        (FP32 accumulation).
// The approximation only screens: a pair is a candidate when
//     approx < t^2 + a (|x|^2 + |y|^2) + b,
// with a and b bounding every error term (below), and each candidate is
// recomputed in the reference's direct FP32 form (fma(diff, diff, acc) over
// k = 0..d-1, the same chain as fph_edge_screen) and kept iff that value
// is < t^2.  So the output -- pairs, dsq bits -- is exactly fph_edge_screen's.
//
// Error bound (u = 2^-24, bf16 rounding to nearest even, unit 2^-8):
//   |x - xh| <= 2^-8 |x|;  x - xh is exact in FP32 (Sterbenz);
//   |(x - xh) - xl| <= 2^-16 |x|, so x = xh + xl + r with |r| <= 2^-16 |x|.
//   Dropped terms per coordinate (xl yl, (xh + xl) ry, rx (yh + yl), rx ry)
//   <= 3.1 * 2^-16 |x_k| |y_k|; bf16 x bf16 products are exact in FP32 and
//   the 3 dpad of them accumulate with error <= gamma(3 dpad) * 1.02 sum
//   |x_k y_k| whatever the accumulation order.  With sum |x_k y_k| <=
//   |x| |y| <= (|x|^2 + |y|^2) / 2:  |2 (x.y - G)| <= cG (|x|^2 + |y|^2),
//   cG = 3.1 * 2^-16 + 1.02 gamma(3 dpad).
//   The FP32 norms add gamma(d) (|x|^2 + |y|^2), the FP32 operations of the
//   test below <= 6u (|x|^2 + |y|^2), and the direct form differs from the
//   true distance by <= gamma(d) dsq ~ gamma(d) t^2 near the threshold.
//   All of it runs on centred points x' = fl(x - c) (c = the coordinate
//   mean; see k_centre), whose rounding adds 2.1u to a and 2u t^2 to b:
//   a = 1.25 (cG + gamma(d) + 8.1u),  b = 1.1 (gamma(d) + 2u) t^2.
// The test is evaluated as fma(-2, G, (1-a)|x|^2 + (1-a)|y|^2) < t^2 + b
// (the scaled norms staged once per tile): three operations per pair.
#include "fph_common.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kMT = 128;       // points per tile side
constexpr int kMTLog = 7;
constexpr int kThreadsM = 256; // 4 waves, each a 64 x 64 quarter of the tile
constexpr int kStageSteps = 2; // 16-dimension k-steps staged in LDS per chunk
static_assert(kMT == 1 << kMTLog, "tile side is a power of two");
constexpr int kSuperLog = 4;   // tiles are taken in 16 x 16 super-tiles
// A chunk holds 2 or 4 vectors of 8 coordinates per point (dpad is a
// multiple of 16), so a point's hi and lo vectors of one chunk are 4 or 8
// contiguous bf16x8: a power of two, decoded with shifts.
static_assert(kStageSteps == 2, "chunk decode assumes 2 or 4 vectors per point");

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

__device__ __forceinline__ unsigned short bf16_rne(float x) {
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
    const unsigned u = (unsigned)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// One thread per (point, coordinate) of the padded coordinates (zero past
// d).  Layout of xhl (bf16, 2 n dpad): per point, per vector of 8
// coordinates, the hi vector then the lo vector, so one point's operands for
// 4 k-steps of 16 coordinates are one 128-byte line.
__global__ void k_split_bf16(const float* __restrict__ pts, const float* __restrict__ ctr,
                             long long n, int d, int dpad, unsigned short* __restrict__ xhl) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * dpad) return;
    const long long i = t / dpad;
    const int k = (int)(t - i * dpad);
    const float x = k < d ? pts[i * d + k] - ctr[k] : 0.0f;
    const unsigned short h = bf16_rne(x);
    const float r = x - bf16_to_f32(h);  // exact
    const long long at = 2 * (i * dpad + (k & ~7)) + (k & 7);
    xhl[at] = h;
    xhl[at + 8] = bf16_rne(r);
}

__global__ void k_norms(const float* __restrict__ pts, const float* __restrict__ ctr,
                        long long n, int d, float* __restrict__ nrm) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    // 8 coordinates' loads in flight together; the FMA chain keeps its order
#pragma unroll 8
    for (int k = 0; k < d; k++) {
        const float x = pts[i * d + k] - ctr[k];
        acc = __builtin_fmaf(x, x, acc);
    }
    nrm[i] = acc;
}

// Centre of the cloud, for the screen only: distances do not change under
// translation, but the margin grows with |x|^2 + |y|^2, so the Gram
// products are formed on x - c with c the coordinate mean (any c keeps the
// bound valid; the mean keeps the norms, hence the margin band, small for a
// cloud far from the origin).  Partial sums per block of rows in FP64, then
// a fixed-order sum per coordinate: deterministic.
constexpr int kCtrBlocks = 256;

__global__ void k_centre_partial(const float* __restrict__ pts, long long n, int d,
                                 double* __restrict__ part) {
    const long long rows = (n + gridDim.x - 1) / gridDim.x;
    const long long r0 = (long long)blockIdx.x * rows;
    const long long r1 = r0 + rows < n ? r0 + rows : n;
    for (int k = threadIdx.x; k < d; k += blockDim.x) {
        double acc = 0.0;
        for (long long i = r0; i < r1; i++) acc += (double)pts[i * d + k];
        part[(long long)blockIdx.x * d + k] = acc;
    }
}

__global__ void k_centre(const double* __restrict__ part, int nb, long long n, int d,
                         float* __restrict__ ctr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d) return;
    double acc = 0.0;
    for (int b = 0; b < nb; b++) acc += part[(long long)b * d + k];
    ctr[k] = (float)(acc / (double)n);
}

__device__ __forceinline__ void tile_coords(long long t, int* I, int* J) {
    long long j = (long long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= t) j++;
    while (j * (j + 1) / 2 > t) j--;
    *J = (int)j;
    *I = (int)(t - j * (j + 1) / 2);
}

// Stages one chunk of the tile's operands: per point, 2^LGP consecutive
// vectors (hi, lo of the chunk's vectors).  Thread tid takes vector slot
// tid % 2^LGP of points tid / 2^LGP + P u (P = 256 / 2^LGP), so consecutive
// lanes read consecutive 16 bytes of one point's line and the LDS addresses
// differ by constant offsets.  Loads go out in groups of kStageGroup with
// constant trip counts: a group's loads are all in flight before the first
// LDS write waits (one memory latency per group, not per load).  Point
// indices pass through an empty asm so that they are recomputed per chunk
// rather than held across the k-loop beside the accumulators.
constexpr int kStageGroup = 8;

template <int LGP>
__device__ __forceinline__ void stage_chunk(const bf16x8* __restrict__ xhl, int I, int J, int n,
                                            int vrow, int kc, int tid,
                                            bf16x8 (*sA)[kMT + 1], bf16x8 (*sB)[kMT + 1]) {
    constexpr int kPts = kThreadsM >> LGP;   // points per pass
    constexpr int kPasses = kMT / kPts;      // passes per side
    constexpr int kLoads = 2 * kPasses;      // both sides
    constexpr int kGroup = kLoads < kStageGroup ? kLoads : kStageGroup;
    static_assert(kPasses * kPts == kMT && kLoads % kGroup == 0, "whole groups");
    const int jv = tid & ((1 << LGP) - 1), p0 = tid >> LGP;
    const bf16x8* src = xhl + 2 * kc + jv;
    const long long stride = 2LL * vrow;
#pragma unroll
    for (int f0 = 0; f0 < kLoads; f0 += kGroup) {
        bf16x8 v[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int f = f0 + u;
            int g = min((f < kPasses ? I : J) * kMT + p0 + (f % kPasses) * kPts, n - 1);
            asm volatile("" : "+v"(g));
            v[u] = src[(long long)g * stride];
        }
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int f = f0 + u;
            bf16x8(*dst)[kMT + 1] = f < kPasses ? sA : sB;
            dst[jv][p0 + (f % kPasses) * kPts] = v[u];
        }
    }
}

__global__ __launch_bounds__(kThreadsM, 4) void k_edge_screen_mfma(
    const float* __restrict__ pts, const bf16x8* __restrict__ xhl,
    const float* __restrict__ nrm, int n, int d, int dpad,
    float thr, float marg_a, float marg_b, long long nslots, unsigned long long* __restrict__ keys,
    unsigned long long* __restrict__ vals, long long cap,
    unsigned long long* __restrict__ count, unsigned long long* __restrict__ ncand,
    unsigned long long* __restrict__ band, long long band_cap,
    unsigned long long* __restrict__ nband) {
    __shared__ float s_nr[kMT], s_nc[kMT];
    // operands of kStageSteps k-steps of the whole tile, [2 vector + h/l]
    // [point]: each row and column block is read from global memory once per
    // tile, not once per wave that uses it.  Rows are padded by one vector:
    // the 8 lanes that stage one point's line write 8 rows at distinct
    // banks, and 8 lanes reading consecutive points of one row do too.
    __shared__ bf16x8 sA[4 * kStageSteps][kMT + 1];
    __shared__ bf16x8 sB[4 * kStageSteps][kMT + 1];
    // XCD-aware order: the 8 XCDs take blockIdx round robin; each XCD takes
    // a contiguous run of slots.  Slots enumerate the 16 x 16 super-tiles of
    // the upper triangle, then their tiles: the tiles of one super-tile run
    // together on one XCD and share 16 row and 16 column blocks in its L2
    // (a plain triangular order re-reads a row block for almost every tile).
    const long long nb = gridDim.x;
    const long long per_xcd = (nb + 7) / 8;
    const long long t = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (t >= nslots) return;  // the whole workgroup, before any barrier
    int SI, SJ;
    tile_coords(t >> (2 * kSuperLog), &SI, &SJ);
    const int l = (int)(t & ((1 << (2 * kSuperLog)) - 1));
    const int I = (SI << kSuperLog) + (l & ((1 << kSuperLog) - 1));
    const int J = (SJ << kSuperLog) + (l >> kSuperLog);
    // below the diagonal or past the last tile: the whole workgroup leaves
    if (I > J || J >= (n + kMT - 1) / kMT) return;
    const int tid = threadIdx.x;
    // this thread's scaled norm: read here, written to LDS with the first
    // chunk's operands, so both loads are in flight together
    const float one_a = 1.0f - marg_a;
    const float my_nrm =
        one_a * nrm[min(tid < kMT ? I * kMT + tid : J * kMT + tid - kMT, n - 1)];
    const int w = tid / FPH_WAVE, lane = fph_lane();
    const int rsub = (w >> 1) * 64, csub = (w & 1) * 64;
    // entirely below the diagonal: stages operands with the others, then leaves
    const bool below = I == J && rsub > csub;
    const int row0 = I * kMT + rsub, col0 = J * kMT + csub;
    const int vrow = dpad / 8;  // bf16x8 vectors per point
    const int r = lane & 31, kh = lane >> 5;
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][b][v] = 0.0f;
    for (int kc = 0; kc < vrow; kc += 2 * kStageSteps) {
        const int nv = min(2 * kStageSteps, vrow - kc);  // vectors in this chunk (even)
        __syncthreads();  // the previous chunk's readers are done
        if (nv == 4)
            stage_chunk<3>(xhl, I, J, n, vrow, kc, tid, sA, sB);
        else
            stage_chunk<2>(xhl, I, J, n, vrow, kc, tid, sA, sB);
        if (kc == 0) {  // the loop runs at least once: vrow >= 2
            if (tid < kMT)
                s_nr[tid] = my_nrm;
            else
                s_nc[tid - kMT] = my_nrm;
        }
        __syncthreads();
        if (below) continue;
        for (int ks = 0; ks < nv / 2; ks++) {
            const int o = 2 * (ks * 2 + kh);  // hi row of this lane's vector; lo at o + 1
            const bf16x8 ah0 = sA[o][rsub + r], al0 = sA[o + 1][rsub + r];
            const bf16x8 ah1 = sA[o][rsub + 32 + r], al1 = sA[o + 1][rsub + 32 + r];
            const bf16x8 bh0 = sB[o][csub + r], bl0 = sB[o + 1][csub + r];
            const bf16x8 bh1 = sB[o][csub + 32 + r], bl1 = sB[o + 1][csub + 32 + r];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, bh0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, bh1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, bh0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, bh1, acc[1][1], 0, 0, 0);
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, bl0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, bl1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, bl0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah1, bl1, acc[1][1], 0, 0, 0);
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0, bh0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0, bh1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1, bh0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al1, bh1, acc[1][1], 0, 0, 0);
        }
    }
    if (below) return;
    // Epilogue: accumulator slot v of lane l is row 8 (v / 4) + 4 (l / 32)
    // + v % 4, column l % 32 of its 32 x 32 block.  Off the diagonal and
    // inside [0, n) every pair is valid (i < j < n).
    const float lim = thr + marg_b;
    const bool interior = I < J && (J + 1) * kMT <= n;
    const float cn0 = s_nc[csub + r], cn1 = s_nc[csub + 32 + r];
    // Candidates are rare: a first pass takes the minimum of the tested
    // values (any value < lim makes the minimum < lim; NaN values, never
    // hits, are skipped by fminf), and a wave with no hit leaves.  The second
    // pass evaluates the same expression per pair.  The running minimum is
    // materialised every 8 pairs so that the 64 values are never live at
    // once beside the accumulators (register budget).
    float vmin = __builtin_inff();
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const float rv = s_nr[rsub + 32 * a + 8 * (v >> 2) + 4 * kh + (v & 3)];
#pragma unroll
            for (int b = 0; b < 2; b++)
                vmin = fminf(vmin, __builtin_fmaf(-2.0f, acc[a][b][v], rv + (b ? cn1 : cn0)));
            if ((v & 3) == 3) asm volatile("" : "+v"(vmin));
        }
    if (__ballot(vmin < lim) == 0ull) return;
    // an opaque copy of the row offset: the second pass reloads the norms
    // instead of keeping the first pass's values live (register budget)
    int rsub2 = rsub;
    asm volatile("" : "+v"(rsub2));
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const float rv = s_nr[rsub2 + 32 * a + 8 * (v >> 2) + 4 * kh + (v & 3)];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const bool hit = __builtin_fmaf(-2.0f, acc[a][b][v], rv + (b ? cn1 : cn0)) < lim;
                const int li = rsub + 32 * a + 8 * (v >> 2) + 4 * kh + (v & 3);
                const int lj = csub + 32 * b + r;
                const int i = I * kMT + li, j = J * kMT + lj;
                const bool cand = hit && (interior || (i < j && j < n));
                if (cand) {  // rare: recompute in the direct form
                    atomicAdd(ncand, 1ull);
                    const float* xi = pts + (long long)i * d;
                    const float* xj = pts + (long long)j * d;
                    float e = 0.0f;
                    // unrolled so that 8 coordinates' loads are in flight
                    // together; the FMA chain keeps its order k = 0..d-1
#pragma unroll 8
                    for (int k = 0; k < d; k++) {
                        const float diff = xi[k] - xj[k];
                        e = __builtin_fmaf(diff, diff, e);
                    }
                    if (e < thr) {
                        const unsigned long long at = atomicAdd(count, 1ull);
                        if (at < (unsigned long long)cap) {
                            unsigned bits;
                            memcpy(&bits, &e, 4);
                            const unsigned long long uj = (unsigned long long)j;
                            const unsigned long long colex = uj * (uj - 1) / 2 + (unsigned)i;
                            keys[at] = ((unsigned long long)bits << 32) | (colex & 0xFFFFFFFFull);
                            if (vals) vals[at] = (uj << 32) | (unsigned long long)i;
                        }
                    } else if (band) {  // validation mode: list the margin band
                        const unsigned long long at = atomicAdd(nband, 1ull);
                        if (at < (unsigned long long)band_cap)
                            band[at] = ((unsigned long long)j << 32) | (unsigned long long)i;
                    }
                }
            }
        }
}

__global__ void k_zero2(unsigned long long* a, unsigned long long* b, unsigned long long* c) {
    if (threadIdx.x == 0) {
        a[0] = 0ull;
        b[0] = 0ull;
        if (c) c[0] = 0ull;
    }
}

// FP64 check of screened pairs (validation mode).  Per pair (val = j << 32 |
// i): the FP32 FMA chain of the filtration and the FP64 distance (FP32
// coordinates differ exactly in FP64; the FP64 FMA chain is within d 2^-53
// relative of |x - y|^2).  st[0] pairs, [1] FP32 members (dsq32 < thr_sq),
// [2] FP64 members (dsq64 < t^2 with t the caller's double), [3] FP32 in /
// FP64 out, [4] FP32 out / FP64 in, [5] largest |dsq32 - dsq64| over FP32
// members in 1/1024 ulp of dsq32.
__global__ void k_pairs_fp64(const float* __restrict__ pts, int d,
                             const unsigned long long* __restrict__ pv, long long m,
                             float thr_sq, double t2, unsigned long long* __restrict__ st) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long f32in = 0, f64in = 0, a = 0, b = 0, err = 0;
    const bool live = e < m;
    if (live) {
        const unsigned long long v = pv[e];
        const long long i = (long long)(v & 0xFFFFFFFFull), j = (long long)(v >> 32);
        const float* xi = pts + i * d;
        const float* xj = pts + j * d;
        float e32 = 0.0f;
        double e64 = 0.0;
        for (int k = 0; k < d; k++) {
            const float diff = xi[k] - xj[k];
            e32 = __builtin_fmaf(diff, diff, e32);
            const double dd = (double)xi[k] - (double)xj[k];
            e64 = __builtin_fma(dd, dd, e64);
        }
        const bool in32 = e32 < thr_sq, in64 = e64 < t2;
        f32in = in32;
        f64in = in64;
        a = in32 && !in64;
        b = !in32 && in64;
        if (in32 && e32 > 0.0f) {
            unsigned bits;
            memcpy(&bits, &e32, 4);
            const unsigned nb = bits + 1u;
            float nxt;
            memcpy(&nxt, &nb, 4);
            const double ulp = (double)nxt - (double)e32;
            err = (unsigned long long)(fabs((double)e32 - e64) / ulp * 1024.0);
        }
    }
    // wave totals, one atomic per wave and counter
    const unsigned long long m_live = __ballot(live);
    const unsigned long long c1 = __popcll(__ballot(f32in)), c2 = __popcll(__ballot(f64in));
    const unsigned long long c3 = __popcll(__ballot(a)), c4 = __popcll(__ballot(b));
    unsigned long long mx = err;
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(mx, o);
        mx = y > mx ? y : mx;
    }
    if (fph_lane() == 0 && m_live) {
        atomicAdd(&st[0], (unsigned long long)__popcll(m_live));
        if (c1) atomicAdd(&st[1], c1);
        if (c2) atomicAdd(&st[2], c2);
        if (c3) atomicAdd(&st[3], c3);
        if (c4) atomicAdd(&st[4], c4);
        if (mx) atomicMax(&st[5], mx);
    }
}

__global__ void k_zero_n(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

double gamma_n(int m) {
    const double u = 1.0 / 16777216.0;
    return m * u / (1.0 - m * u);
}

}  // namespace

static int screen_mfma(const float* pts, int64_t n, int d, float thr_sq, uint64_t* keys,
                       uint64_t* vals, int64_t cap, uint64_t* count_dev, uint64_t* cand_dev,
                       uint64_t* band, int64_t band_cap, uint64_t* band_count, void* ws,
                       size_t* ws_bytes, hipStream_t s) {
    if (n < 0 || d < 1 || cap < 0 || !count_dev || !cand_dev || !ws_bytes) return FPH_ERR_ARG;
    if (band && (band_cap < 0 || !band_count)) return FPH_ERR_ARG;
    if (n > 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    const int dpad = (d + 15) / 16 * 16;
    const size_t split = align_up(4 * (size_t)std::max<int64_t>(n, 1) * dpad);
    const size_t nrm_b = align_up(4 * (size_t)std::max<int64_t>(n, 1));
    const size_t ctr_b = align_up(4 * (size_t)d);
    const size_t part_b = align_up(8 * (size_t)kCtrBlocks * d);
    const size_t need = split + nrm_b + ctr_b + part_b;
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    unsigned short* xhl = (unsigned short*)w;
    float* nrm = (float*)(w + split);
    float* ctr = (float*)(w + split + nrm_b);
    double* part = (double*)(w + split + nrm_b + ctr_b);
    k_zero2<<<1, 64, 0, s>>>((unsigned long long*)count_dev, (unsigned long long*)cand_dev,
                             band ? (unsigned long long*)band_count : nullptr);
    FPH_LAUNCH_CHECK();
    if (n < 2) return FPH_OK;
    k_centre_partial<<<kCtrBlocks, 64, 0, s>>>(pts, n, d, part);
    k_centre<<<fph_blocks(d, 64), 64, 0, s>>>(part, kCtrBlocks, n, d, ctr);
    k_split_bf16<<<fph_blocks(n * dpad, 256), 256, 0, s>>>(pts, ctr, n, d, dpad, xhl);
    FPH_LAUNCH_CHECK();
    k_norms<<<fph_blocks(n, 256), 256, 0, s>>>(pts, ctr, n, d, nrm);
    FPH_LAUNCH_CHECK();
    // The bound of the header, on x' = fl(x - c): |x'_k - (x_k - c_k)| <=
    // u' |x'_k| (u' = u / (1 - u)), so | |x'-y'|^2 - |x-y|^2 | <= u' |x-y|^2 +
    // (2u' + 2u'^2)(|x'|^2 + |y'|^2): 2.1u more in a, and |x-y|^2's factor
    // (1 + u') near t^2 adds u t^2 to b (2u t^2 is used).
    const double cg = 3.1 * std::ldexp(1.0, -16) + 1.02 * gamma_n(3 * dpad);
    const double u = std::ldexp(1.0, -24);
    const float marg_a = (float)(1.25 * (cg + gamma_n(d) + 8.1 * u));
    const float marg_b = (float)(1.1 * (gamma_n(d) + 2.0 * u) * (double)thr_sq);
    const long long nts = (n + kMT - 1) / kMT;
    const long long nsup = (nts + (1 << kSuperLog) - 1) >> kSuperLog;
    const long long nslots = (nsup * (nsup + 1) / 2) << (2 * kSuperLog);
    if (nslots > 0x7FFFFFF0LL) return FPH_ERR_LIMIT;
    // 8 x ceil(nslots / 8) blocks: the XCD remap is a bijection onto
    // [0, 8 ceil(nslots / 8)); blocks past nslots exit.
    const long long nblk = (nslots + 7) / 8 * 8;
    k_edge_screen_mfma<<<(unsigned)nblk, kThreadsM, 0, s>>>(
        pts, (const bf16x8*)xhl, nrm, (int)n, d, dpad, thr_sq, marg_a, marg_b,
        nslots, (unsigned long long*)keys, (unsigned long long*)vals, cap,
        (unsigned long long*)count_dev, (unsigned long long*)cand_dev,
        (unsigned long long*)band, band_cap, (unsigned long long*)band_count);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_edge_screen_mfma(const float* pts, int64_t n, int d, float thr_sq,
                                    uint64_t* keys, uint64_t* vals, int64_t cap,
                                    uint64_t* count_dev, uint64_t* cand_dev, void* ws,
                                    size_t* ws_bytes, void* stream) {
    return screen_mfma(pts, n, d, thr_sq, keys, vals, cap, count_dev, cand_dev, nullptr, 0,
                       nullptr, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int fph_edge_screen_mfma_band(const float* pts, int64_t n, int d, float thr_sq,
                                         uint64_t* keys, uint64_t* vals, int64_t cap,
                                         uint64_t* count_dev, uint64_t* cand_dev,
                                         uint64_t* band, int64_t band_cap,
                                         uint64_t* band_count, void* ws, size_t* ws_bytes,
                                         void* stream) {
    if (!band) return FPH_ERR_ARG;
    return screen_mfma(pts, n, d, thr_sq, keys, vals, cap, count_dev, cand_dev, band, band_cap,
                       band_count, ws, ws_bytes, (hipStream_t)stream);
}

extern "C" int fph_pairs_fp64_check(const float* pts, int64_t n, int d, double max_edge_length,
                                    float thr_sq, const uint64_t* pair_vals, int64_t m,
                                    uint64_t* stats, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (n < 0 || d < 1 || m < 0 || !stats || (m > 0 && !pair_vals)) return FPH_ERR_ARG;
    k_zero_n<<<1, 64, 0, s>>>((unsigned long long*)stats, 8);
    FPH_LAUNCH_CHECK();
    if (m == 0) return FPH_OK;
    k_pairs_fp64<<<fph_blocks(m, 256), 256, 0, s>>>(
        pts, d, (const unsigned long long*)pair_vals, m, thr_sq,
        max_edge_length * max_edge_length, (unsigned long long*)stats);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
