// Stage (e): residual Z/2 cohomology reduction on the GPU.
//
// Replaces the Numba loop of cohomology_general.py:202-371 (with the pool
// prefill of _numba_utils.py:68-104 and the hash map of _numba_utils.py:14-40).
//
// Columns (edges for H1, triangles for H2) are processed by the reference in
// descending rank order; a column's pivot is its smallest row; a column whose
// pivot is owned by an earlier (higher-rank) column is XORed with that
// column.  Every reduced matrix obtained by adding earlier columns to later
// ones has the same pivot pairs, so the same pairs are obtained here without
// the serial order:
//
//   phase A  (all residual columns at once, one wavefront per column)
//            XOR with apparent columns while the pivot is owned by one.
//            Apparent columns own their pivot forever: no column of higher
//            rank ever contains it.
//   phase B  (one persistent workgroup, rounds separated by barriers)
//            each column claims its pivot with atomicMax(own[pivot], column);
//            the highest-rank claimant owns it and every other claimant (and
//            any displaced previous owner) is XORed with the owner in the
//            next round, then chains apparent XORs again.  Owners never
//            change inside a round, so reads and writes never overlap.
//            Stops when no two nonzero columns share a pivot.
//
// A column is a sorted list of row ranks.  The XOR of two sorted lists is a
// merge with cancellation, computed by a whole wavefront: every element
// binary-searches the other list; for a kept element a = A[i],
//     position = 2 * (#kept A before i) + lower_bound(B, a) - i,
// and symmetrically for B.  Lists live in LDS when they fit, in a
// bump-allocated global pool otherwise.
#include "fph_common.h"

namespace {

constexpr int kCapW = 320;  // LDS words per working buffer (two per wave)
constexpr int kCapO = 256;  // LDS words for the staged owner column
constexpr int kWaveLds = 2 * kCapW + kCapO;
constexpr int kPhaseAWaves = 4;
constexpr int kPhaseBWaves = 16;

struct RedArgs {
    const int* off;                 // CSR row offsets, K + 1
    const unsigned* data;           // CSR rows, ascending global row ranks
    long long row_base;             // global rank of row 0
    const unsigned char* kind;      // 2 = apparent, 3 = residual
    int* own;                       // nrows: owning column of a pivot row or -1
    int* pivot;                     // K: current pivot (global rank) or -1
    unsigned long long* col_ptr;    // K: address of the current column content
    int* col_len;                   // K
    unsigned* pool;
    unsigned long long pool_cap;
    unsigned long long* pool_ptr;
    int* status;                    // [0] pool overflow, [1] rounds, [2] xors
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int lb_u32(const unsigned* x, int n, unsigned v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// C = A xor B for sorted duplicate-free lists; C must not alias A or B.
__device__ int wave_xor(const unsigned* A, int na, const unsigned* B, int nb, unsigned* C) {
    const int lane = fph_lane();
    const unsigned long long lt = fph_lanemask_lt();
    int run_a = 0;
    for (int base = 0; base < na; base += FPH_WAVE) {
        const int i = base + lane;
        bool keep = false;
        unsigned a = 0;
        int pos = 0;
        if (i < na) {
            a = A[i];
            const int lb = lb_u32(B, nb, a);
            keep = !(lb < nb && B[lb] == a);
            pos = lb - i;
        }
        const unsigned long long bal = __ballot(keep);
        if (keep) C[2 * (run_a + __popcll(bal & lt)) + pos] = a;
        run_a += __popcll(bal);
    }
    int run_b = 0;
    for (int base = 0; base < nb; base += FPH_WAVE) {
        const int j = base + lane;
        bool keep = false;
        unsigned b = 0;
        int pos = 0;
        if (j < nb) {
            b = B[j];
            const int lb = lb_u32(A, na, b);
            keep = !(lb < na && A[lb] == b);
            pos = lb - j;
        }
        const unsigned long long bal = __ballot(keep);
        if (keep) C[2 * (run_b + __popcll(bal & lt)) + pos] = b;
        run_b += __popcll(bal);
    }
    wave_sync();
    return run_a + run_b;
}

__device__ __forceinline__ void wave_copy(const unsigned* src, int n, unsigned* dst) {
    for (int i = fph_lane(); i < n; i += FPH_WAVE) dst[i] = src[i];
    wave_sync();
}

// Wave-uniform pool allocation; returns nullptr on overflow (flag raised).
__device__ unsigned* pool_alloc(const RedArgs& a, int n) {
    unsigned long long at = 0;
    if (fph_lane() == 0) at = atomicAdd(a.pool_ptr, (unsigned long long)n);
    at = __shfl(at, 0);
    if (at + (unsigned long long)n > a.pool_cap) {
        if (fph_lane() == 0) atomicExch(&a.status[0], 1);
        return nullptr;
    }
    return a.pool + at;
}

struct Work {
    const unsigned* ptr;
    int len;
    int where;  // 0 external (read-only), 1 LDS buf0, 2 LDS buf1, 3 pool
};

__device__ __forceinline__ void work_load(Work& w, const unsigned* src, int n, unsigned* lds) {
    if (n <= kCapW) {
        wave_copy(src, n, lds);
        w.ptr = lds;
        w.where = 1;
    } else {
        w.ptr = src;
        w.where = 0;
    }
    w.len = n;
}

// w <- w xor B.  Returns false on pool overflow.
__device__ bool xor_step(Work& w, const unsigned* B, int nb, unsigned* lds, const RedArgs& a) {
    unsigned* b0 = lds;
    unsigned* b1 = lds + kCapW;
    unsigned* st = lds + 2 * kCapW;
    const int need = w.len + nb;
    unsigned* out;
    int where;
    if (need <= kCapW) {
        out = (w.where == 1) ? b1 : b0;
        where = (w.where == 1) ? 2 : 1;
    } else {
        out = pool_alloc(a, need);
        if (!out) return false;
        where = 3;
    }
    const unsigned* bs = B;
    if (nb <= kCapO) {
        wave_copy(B, nb, st);
        bs = st;
    }
    const unsigned p_old = w.ptr[0];
    w.len = wave_xor(w.ptr, w.len, bs, nb, out);
    if (where == 3) __threadfence();
    w.ptr = out;
    w.where = where;
    // The XOR of two columns with the same pivot removes it, so the pivot
    // strictly increases; anything else means inconsistent inputs.
    if (w.len > 0 && !(w.ptr[0] > p_old)) {
        if (fph_lane() == 0) atomicExch(&a.status[0], 3);
        return false;
    }
    return true;
}

// XOR with apparent owners while the pivot has one.
__device__ bool chain_apparent(Work& w, unsigned* lds, const RedArgs& a, int* nxor) {
    while (w.len > 0) {
        const unsigned p = w.ptr[0];
        const int o = a.own[(long long)p - a.row_base];
        if (o < 0 || a.kind[o] != 2) break;
        const int lo = a.off[o];
        if (!xor_step(w, a.data + lo, a.off[o + 1] - lo, lds, a)) return false;
        (*nxor)++;
    }
    return true;
}

// Publish the working column as column c's content.
__device__ bool work_store(Work& w, int c, const unsigned* orig, const RedArgs& a, bool changed) {
    const unsigned* final_ptr = w.ptr;
    if (!changed) {
        final_ptr = orig;
    } else if (w.where == 1 || w.where == 2) {
        unsigned* dst = w.len > 0 ? pool_alloc(a, w.len) : a.pool;
        if (!dst) return false;
        wave_copy(w.ptr, w.len, dst);
        __threadfence();
        final_ptr = dst;
    }
    const unsigned piv = w.len > 0 ? w.ptr[0] : 0u;
    if (fph_lane() == 0) {
        a.col_ptr[c] = (unsigned long long)(uintptr_t)final_ptr;
        a.col_len[c] = w.len;
        a.pivot[c] = w.len > 0 ? (int)piv : -1;
    }
    return true;
}

__global__ __launch_bounds__(kPhaseAWaves * FPH_WAVE) void k_reduce_chain(
    RedArgs a, const int* __restrict__ res_list, const int* __restrict__ res_count) {
    extern __shared__ unsigned smem[];
    const int wid = threadIdx.x / FPH_WAVE;
    unsigned* lds = smem + wid * kWaveLds;
    const int R = *res_count;
    const long long nw = (long long)gridDim.x * kPhaseAWaves;
    int nxor = 0;
    for (long long r = (long long)blockIdx.x * kPhaseAWaves + wid; r < R; r += nw) {
        if (a.status[0]) return;
        const int c = res_list[r];
        const int lo = a.off[c];
        const unsigned* orig = a.data + lo;
        Work w;
        work_load(w, orig, a.off[c + 1] - lo, lds);
        const int before = nxor;
        if (!chain_apparent(w, lds, a, &nxor)) return;
        if (!work_store(w, c, orig, a, nxor != before)) return;
    }
    if (fph_lane() == 0 && nxor) atomicAdd(&a.status[2], nxor);
}

__device__ __forceinline__ void claim(const RedArgs& a, int c, int* wl, int* cnt) {
    const int p = a.pivot[c];
    const int old = atomicMax(&a.own[(long long)p - a.row_base], c);
    if (old > c) wl[atomicAdd(cnt, 1)] = c;
    else if (old >= 0) wl[atomicAdd(cnt, 1)] = old;
}

__global__ __launch_bounds__(kPhaseBWaves * FPH_WAVE) void k_reduce_rounds(
    RedArgs a, const int* __restrict__ res_list, const int* __restrict__ res_count,
    int* __restrict__ wl0, int* __restrict__ wl1, int max_rounds) {
    extern __shared__ unsigned smem[];
    __shared__ int s_cnt[2];
    __shared__ int s_abort;
    const int tid = threadIdx.x;
    const int wid = tid / FPH_WAVE;
    unsigned* lds = smem + wid * kWaveLds;
    const int R = *res_count;
    if (tid == 0) {
        s_cnt[0] = 0;
        s_cnt[1] = 0;
        s_abort = a.status[0];
    }
    __syncthreads();
    if (s_abort) return;
    for (int r = tid; r < R; r += blockDim.x) {
        const int c = res_list[r];
        if (a.col_len[c] > 0) claim(a, c, wl0, &s_cnt[0]);
    }
    __syncthreads();
    int cur = 0, rounds = 0, nxor = 0;
    while (true) {
        const int n = s_cnt[cur];
        if (n == 0 || s_abort || rounds >= max_rounds) break;
        int* wl = cur ? wl1 : wl0;
        int* wn = cur ? wl0 : wl1;
        for (int k = wid; k < n; k += kPhaseBWaves) {
            if (s_abort) break;
            const int c = wl[k];
            const int p = a.pivot[c];
            const int o = a.own[(long long)p - a.row_base];
            if (o <= c) {  // an owner is always a higher-rank column
                if (fph_lane() == 0) {
                    atomicExch(&a.status[0], 3);
                    s_abort = 1;
                }
                break;
            }
            const unsigned* ob;
            int on;
            if (a.kind[o] == 2) {  // apparent owners keep their CSR row
                ob = a.data + a.off[o];
                on = a.off[o + 1] - a.off[o];
            } else {
                ob = (const unsigned*)(uintptr_t)a.col_ptr[o];
                on = a.col_len[o];
            }
            Work w;
            const unsigned* orig = (const unsigned*)(uintptr_t)a.col_ptr[c];
            work_load(w, orig, a.col_len[c], lds);
            bool ok = xor_step(w, ob, on, lds, a);
            nxor++;
            if (ok) ok = chain_apparent(w, lds, a, &nxor);
            if (ok) ok = work_store(w, c, orig, a, true);
            if (!ok) {
                if (fph_lane() == 0) s_abort = 1;
                break;
            }
        }
        __syncthreads();
        if (tid == 0) s_cnt[cur ^ 1] = 0;
        __syncthreads();
        if (!s_abort) {
            for (int k = tid; k < n; k += blockDim.x) {
                const int c = wl[k];
                if (a.col_len[c] > 0) claim(a, c, wn, &s_cnt[cur ^ 1]);
            }
        }
        __syncthreads();
        cur ^= 1;
        rounds++;
    }
    if (tid == 0) {
        a.status[1] = rounds;
        if (s_cnt[cur] != 0 && !s_abort) a.status[0] = 2;  // round cap reached
    }
    if (fph_lane() == 0 && nxor) atomicAdd(&a.status[2], nxor);
}

__global__ void k_reduce_init(int* __restrict__ status, unsigned long long* __restrict__ pool_ptr) {
    if (threadIdx.x == 0) {
        status[0] = 0;
        status[1] = 0;
        status[2] = 0;
        pool_ptr[0] = 0;
    }
}

}  // namespace

extern "C" size_t fph_reduce_scratch_bytes(int64_t K) {
    const size_t k = (size_t)std::max<int64_t>(K, 1);
    // col_ptr (8K) + col_len (4K) + two worklists (8K) + pool_ptr + status
    return 8 * k + 4 * k + 8 * k + 256;
}

extern "C" int fph_reduce(int64_t K, const int32_t* col_off, const uint32_t* col_data,
                          int64_t row_base, const uint8_t* kind, int32_t* own, int32_t* pivot,
                          const int32_t* res_list, const int32_t* res_count, uint32_t* pool,
                          int64_t pool_cap, void* scratch, size_t scratch_bytes,
                          int32_t* status, int max_rounds, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || pool_cap < 0 || !scratch || !status) return FPH_ERR_ARG;
    if (scratch_bytes < fph_reduce_scratch_bytes(K)) return FPH_ERR_ARG;
    const size_t k = (size_t)std::max<int64_t>(K, 1);
    char* w = (char*)scratch;
    RedArgs a;
    a.off = col_off;
    a.data = col_data;
    a.row_base = row_base;
    a.kind = kind;
    a.own = own;
    a.pivot = pivot;
    a.col_ptr = (unsigned long long*)w;
    a.col_len = (int*)(w + 8 * k);
    int* wl0 = (int*)(w + 12 * k);
    int* wl1 = (int*)(w + 16 * k);
    a.pool_ptr = (unsigned long long*)(w + 20 * k + ((8 - (20 * k) % 8) % 8));
    a.pool = pool;
    a.pool_cap = (unsigned long long)pool_cap;
    a.status = status;
    k_reduce_init<<<1, 64, 0, s>>>(status, a.pool_ptr);
    FPH_LAUNCH_CHECK();
    if (K == 0) return FPH_OK;
    const unsigned grid_a = 1024;
    k_reduce_chain<<<grid_a, kPhaseAWaves * FPH_WAVE, kPhaseAWaves * kWaveLds * 4, s>>>(
        a, res_list, res_count);
    FPH_LAUNCH_CHECK();
    k_reduce_rounds<<<1, kPhaseBWaves * FPH_WAVE, kPhaseBWaves * kWaveLds * 4, s>>>(
        a, res_list, res_count, wl0, wl1, max_rounds);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
