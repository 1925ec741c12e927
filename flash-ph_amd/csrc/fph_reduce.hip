// Stage (e): residual Z/2 cohomology reduction on the GPU.
//
// Replaces the Numba loop of cohomology_general.py:202-371 (with the pool
// prefill of _numba_utils.py:68-104 and the hash map of _numba_utils.py:14-40).
//
// Columns (edges for H1, triangles for H2) are reduced by the reference in
// descending rank order: a column's pivot is its smallest row, and a column
// whose pivot is owned by an earlier (higher-rank) column is XORed with that
// column.  Any reduction that only adds higher-rank columns to lower-rank ones
// ends with the same pivot pairs, which lets the columns proceed concurrently:
//
//   own[row] (64 bits) = (owner column + 1) << 32 | pool offset of the owner's
//   content.  A column owns its pivot while own[pivot] names it; an owner's
//   content never changes while it owns, and every content starts with its
//   owner's pivot, so any column of lower rank may XOR with whatever own[]
//   currently names.
//
//   phase A  (every residual column at once, one wavefront each): XOR with
//            apparent owners while the pivot has one.  Apparent columns own
//            their pivot forever (no higher-rank column contains it).
//   phase B  (one workgroup of 16 wavefronts per segment): a queue of the
//            residual columns in descending rank, then of displaced owners.
//            A wavefront takes a column and XORs it with the current owner
//            of its pivot until the pivot is free or owned by a lower-rank
//            column; then it publishes the content and claims the pivot with
//            one 64-bit atomicMax.  A lost race continues with the winner; a
//            displaced lower-rank owner goes back into the queue.  Ends when
//            no column is pending.  With one segment the queue can instead
//            span several workgroups (k_reduce_queue_dev, the default from
//            8192 columns): ordering claims with device-scope fences would
//            write back and invalidate the per-XCD L2s on every claim
//            (measured 20x slower), so there published columns are
//            write-once over a fill value and readers wait, word by word,
//            with device-coherent loads until the column is complete.
//
// A column is a sorted list of row ranks stored in a bump-allocated pool as
// [length, rows...].  XOR of two sorted lists is a merge with cancellation
// computed by a whole wavefront: every element binary-searches the other
// list; a kept a = A[i] lands at 2 * (#kept A before i) + lower_bound(B, a) - i
// (symmetrically for B).  The working column lives in LDS while it fits.
#include "fph_common.h"

#include <rocprim/rocprim.hpp>

namespace {

// Per-wavefront LDS arena: [half 0 | half 1 | stage].  The working column
// lives in one half, an XOR writes the other half (no copy back), and a long
// owner column is staged in the stage area.  Phase B: 16 waves x 9.75 KB.
constexpr int kHalfA = 768, kStageA = 512;    // phase A: 4 waves x 8 KB
constexpr int kHalfB = 1024, kStageB = 448;   // phase B: 16 waves x 9.75 KB
constexpr int kArenaA = 2 * kHalfA + kStageA;
constexpr int kArenaB = 2 * kHalfB + kStageB;
constexpr int kPhaseAWaves = 4;
constexpr int kPhaseBWaves = 16;
constexpr unsigned long long kNoOwner = 0ull;

struct RedArgs {
    const int* off;                 // CSR row offsets, K + 1
    const unsigned* data;           // CSR rows, ascending global row ranks
    long long row_base;             // global rank of row 0
    long long nrows;                // rows (length of own)
    long long K;                    // columns
    const unsigned char* kind;      // 2 = apparent, 3 = residual
    unsigned long long* own;        // nrows: owner encoding (see above), 0 = none
    int* pivot;                     // K: final pivot (global rank) or -1
    unsigned* desc;                 // K: pool offset of each residual column
    unsigned* pool;
    unsigned long long pool_cap;
    unsigned long long* pool_ptr;
    int* status;  // [0] error, [1] ran, [2] XORs, [3] longest column, [4] off-chip XORs, ...
    // owner lines (device-wide bitmap queue): per row kLineWords 64-bit
    // words ((owner + 1) << 32 | value), values = the owner column's length
    // and first rows; see reduce_claim_dev_bm.  use_lines asks for them;
    // `lines` is set in the kernels (lines_of), null when not used.
    int use_lines;
    unsigned long long* lines;
};

constexpr int kLineWords = 16;  // words per owner line: length + 15 rows

// Word offset of the owner lines in the pool: right after the apparent
// region (off[K] + K words, line-rounded) and one line of padding.
__device__ __forceinline__ unsigned long long lines_base(const int* off, long long K) {
    return ((unsigned long long)off[K] + (unsigned long long)K + 31ull) / 32ull * 32ull + 32ull;
}

// The owner lines of this reduction, or null: not asked for, or no room for
// them below the pool capacity (the same test everywhere, so every kernel
// agrees).
__device__ __forceinline__ unsigned long long* lines_of(const RedArgs& a) {
    if (!a.use_lines) return nullptr;
    const unsigned long long base = lines_base(a.off, a.K);
    if (base + 2ull * kLineWords * (unsigned long long)a.nrows > a.pool_cap) return nullptr;
    return (unsigned long long*)(a.pool + base);
}

__device__ __forceinline__ void write_line_word(const RedArgs& a, long long row, int k, int c,
                                                unsigned value) {
    __hip_atomic_store(&a.lines[row * kLineWords + k],
                       ((unsigned long long)(unsigned)(c + 1) << 32) | value, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// status[6..7]: 64-bit count of entries merged (na + nb + result length
// per column addition), for the algorithmic traffic of the reduction.
__device__ __forceinline__ void count_merged(const RedArgs& a, long long n) {
    atomicAdd(reinterpret_cast<unsigned long long*>(a.status + 6), (unsigned long long)n);
}

// Waits until this wave's vector memory operations (loads, stores and
// atomics, with or without a returned value) have completed.
__device__ __forceinline__ void vm_drain() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ unsigned long long load_own(const RedArgs& a, unsigned p) {
    return __hip_atomic_load(&a.own[(long long)p - a.row_base], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
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

// C = A xor B for sorted duplicate-free lists (na, nb >= 1); C must not
// alias A or B.  Only the heads up to min(max A, max B) interleave: with
// sa = na, sb = lower_bound(B, max A + 1) when max A <= max B (or the mirror
// image), every element of a tail sits after all of the other list, so the
// tail moves by a constant (its list's index + the other length - 2 * the
// number of cancelled pairs m).  Head elements find their position by a
// binary search in the other head:
//     pos(A[i]) = 2 * (#kept A before i) + lower_bound(B, A[i]) - i.
__device__ int wave_xor(const unsigned* A, int na, const unsigned* B, int nb, unsigned* C) {
    const int lane = fph_lane();
    const unsigned long long lt = fph_lanemask_lt();
    const unsigned amax = A[na - 1], bmax = B[nb - 1];
    int sa = na, sb = nb;
    if (amax <= bmax) sb = lb_u32(B, nb, amax + 1u);
    else sa = lb_u32(A, na, bmax + 1u);
    int kept_a = 0;
    for (int base = 0; base < sa; base += FPH_WAVE) {
        const int i = base + lane;
        bool keep = false;
        unsigned a = 0;
        int pos = 0;
        if (i < sa) {
            a = A[i];
            const int lb = lb_u32(B, sb, a);
            keep = !(lb < sb && B[lb] == a);
            pos = lb - i;
        }
        const unsigned long long bal = __ballot(keep);
        if (keep) C[2 * (kept_a + __popcll(bal & lt)) + pos] = a;
        kept_a += __popcll(bal);
    }
    const int m = sa - kept_a;  // cancelled pairs (all inside the heads)
    int kept_b = 0;
    for (int base = 0; base < sb; base += FPH_WAVE) {
        const int j = base + lane;
        bool keep = false;
        unsigned b = 0;
        int pos = 0;
        if (j < sb) {
            b = B[j];
            const int lb = lb_u32(A, sa, b);
            keep = !(lb < sa && A[lb] == b);
            pos = lb - j;
        }
        const unsigned long long bal = __ballot(keep);
        if (keep) C[2 * (kept_b + __popcll(bal & lt)) + pos] = b;
        kept_b += __popcll(bal);
    }
    for (int i = sa + lane; i < na; i += FPH_WAVE) C[i + nb - 2 * m] = A[i];
    for (int j = sb + lane; j < nb; j += FPH_WAVE) C[j + na - 2 * m] = B[j];
    wave_sync();
    return na + nb - 2 * m;
}

// C = A xor B for long lists: the merged order is cut into 64 contiguous
// pieces along the merge path (ties: A first), one per lane; a piece never
// separates an element of A from its equal in B.  Each lane walks its piece
// sequentially twice: once to count the survivors (then a wave prefix sum
// gives its output offset), once to write them.  Sequential reads replace
// the per-element binary searches of wave_xor, which cost one dependent
// memory round trip per step when the lists live off chip.
__device__ int wave_xor_long(const unsigned* A, int na, const unsigned* B, int nb, unsigned* C,
                             int limit = 0x7FFFFFFF) {
    const int lane = fph_lane();
    const int total = na + nb;
    const int per = (total + FPH_WAVE - 1) / FPH_WAVE;
    const int k = min(lane * per, total);
    int lo = max(0, k - nb), hi = min(k, na);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] <= B[k - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    int i0 = lo;
    const int j0 = k - lo;
    if (i0 > 0 && j0 < nb && A[i0 - 1] == B[j0]) i0--;
    int i1 = __shfl(i0, (lane + 1) & (FPH_WAVE - 1));
    int j1 = __shfl(j0, (lane + 1) & (FPH_WAVE - 1));
    if (lane == FPH_WAVE - 1) {
        i1 = na;
        j1 = nb;
    }
    constexpr unsigned kEnd = 0xFFFFFFFFu;  // above every row rank
    // branch-free merge steps: the smaller head survives unless the heads
    // are equal; each list advances when its head was <= the other's
    int cnt = 0;
    {
        int ia = i0, ib = j0;
        unsigned av = ia < i1 ? A[ia] : kEnd, bv = ib < j1 ? B[ib] : kEnd;
        while (av != kEnd || bv != kEnd) {
            cnt += av != bv;
            ia += av <= bv;
            ib += bv <= av;
            av = ia < i1 ? A[ia] : kEnd;
            bv = ib < j1 ? B[ib] : kEnd;
        }
    }
    int incl = cnt;
    for (int o = 1; o < FPH_WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const int len = min(__shfl(incl, FPH_WAVE - 1), limit);
    int at = incl - cnt;
    if (at < len) {
        // with a limit, only the first `limit` survivors are kept (a prefix
        // of the sorted result)
        int ia = i0, ib = j0;
        unsigned av = ia < i1 ? A[ia] : kEnd, bv = ib < j1 ? B[ib] : kEnd;
        while ((av != kEnd || bv != kEnd) && at < len) {
            if (av != bv) C[at++] = min(av, bv);
            ia += av <= bv;
            ib += bv <= av;
            av = ia < i1 ? A[ia] : kEnd;
            bv = ib < j1 ? B[ib] : kEnd;
        }
    }
    wave_sync();
    return len;
}

__device__ __forceinline__ void wave_copy(const unsigned* src, int n, unsigned* dst) {
    for (int i = fph_lane(); i < n; i += FPH_WAVE) dst[i] = src[i];
    wave_sync();
}

// Pool allocations are rounded up to whole 128-byte lines (32 words), so a
// speculative one-line read at the start of a column never touches another
// column's words.  Each wavefront carves its allocations out of a private
// chunk and refills the chunk with one atomic.
constexpr int kLine = 32;

struct Alloc {
    unsigned long long cur;
    unsigned long long end;
    unsigned long long chunk;  // words reserved per refill
    // Two off-chip working buffers per wavefront (pool offsets, capacities
    // in words), reused by every long XOR of the wavefront; grown by
    // doubling when a longer result comes along.
    unsigned long long big_off[3];
    int big_cap[3];  // [2]: staging of long owner columns (device-wide queue)
    // statistics kept per wavefront and added once at the end (device-wide
    // queue: a shared counter updated per addition would serialise them)
    unsigned long long merged;
    int offchip;
    int maxlen;
};

__device__ __forceinline__ Alloc make_alloc(unsigned long long chunk) {
    Alloc al;
    al.cur = al.end = 0;
    al.chunk = chunk;
    al.big_off[0] = al.big_off[1] = al.big_off[2] = 0;
    al.big_cap[0] = al.big_cap[1] = al.big_cap[2] = 0;
    al.merged = 0;
    al.offchip = 0;
    al.maxlen = 0;
    return al;
}

// Wave-uniform allocation of n words; returns the word offset, or -1 on
// overflow (error flag raised).
__device__ long long pool_alloc(const RedArgs& a, Alloc& al, int n) {
    const unsigned long long need = ((unsigned long long)n + kLine - 1) / kLine * kLine;
    if (al.cur + need > al.end) {
        const unsigned long long take = need > al.chunk ? need : al.chunk;
        unsigned long long at = 0;
        if (fph_lane() == 0) at = atomicAdd(a.pool_ptr, take);
        at = __shfl(at, 0);
        if (at + take > a.pool_cap || at + take > 0xFFFFFFFFull) {
            if (fph_lane() == 0) atomicExch(&a.status[0], 1);
            return -1;
        }
        al.cur = at;
        al.end = at + take;
    }
    const unsigned long long at = al.cur;
    al.cur += need;
    return (long long)at;
}

// Speculative read of the first line of a pooled column [len, rows...]:
// lane l < 32 holds word l; returns len (one round trip for short columns).
__device__ __forceinline__ int read_line(const unsigned* src, unsigned& v) {
    v = fph_lane() < kLine ? src[fph_lane()] : 0u;
    return (int)__shfl(v, 0);
}

// Finish copying the rows of a column whose first line is in v.
__device__ __forceinline__ void finish_copy(const unsigned* src, unsigned v, int len,
                                            unsigned* dst) {
    const int lane = fph_lane();
    if (lane >= 1 && lane < kLine && lane <= len) dst[lane - 1] = v;
    for (int i = kLine + lane; i <= len; i += FPH_WAVE) dst[i - 1] = src[i];
    wave_sync();
}

struct Work {
    const unsigned* ptr;  // current rows (arena half, pool or CSR, read-only)
    int len;
    int half;             // arena half holding the rows, -1 when off chip
};

// Start from the pooled column [len, rows...] at src: on chip (half 0) when
// it fits, else read in place.
__device__ __forceinline__ void work_load(Work& w, const unsigned* src, unsigned* arena, int H) {
    unsigned v;
    const int len = read_line(src, v);
    w.len = len;
    if (len <= H) {
        finish_copy(src, v, len, arena);
        w.ptr = arena;
        w.half = 0;
    } else {
        w.ptr = src + 1;
        w.half = -1;
    }
}

// Off-chip working buffer i of the wavefront with room for n words, or
// nullptr on pool overflow.
__device__ unsigned* big_buffer(const RedArgs& a, Alloc& al, int i, int n) {
    if (al.big_cap[i] < n) {
        int cap = 4096;
        while (cap < n) cap *= 2;
        const long long at = pool_alloc(a, al, cap);
        if (at < 0) return nullptr;
        al.big_off[i] = (unsigned long long)at;
        al.big_cap[i] = cap;
    }
    return a.pool + al.big_off[i];
}

// ---- device-coherent access for the device-wide queue --------------------
// Columns published during that queue are written once into pool space
// filled beforehand with kUnwritten (no row rank or length takes that
// value).  Readers use device-coherent loads and wait until no word of
// [len, rows...] reads kUnwritten, so publication needs no fence: each word
// is coherent on its own, and a column is complete once every word is.
constexpr unsigned kUnwritten = 0xFFFFFFFFu;

__device__ __forceinline__ unsigned ld_coh(const unsigned* p) {
    return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup-scope load (misses the CU's L1 but may hit the L2).  Under the
// write-once rule any value other than kUnwritten it returns is final; a
// kUnwritten one is re-read with ld_coh.
__device__ __forceinline__ unsigned ld_fast(const unsigned* p) {
    return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void st_coh(unsigned* p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait whose clock starts at the first retry: the fast path (the
// data already there) reads no clock.  Reading the device real-time clock
// is a scalar-memory operation, and LDS operations wait on the same counter,
// so a read on every column addition stalls the wavefront's next LDS access
// for as long as the clock read takes -- long when hundreds of polling
// wavefronts read the clock too.
struct Deadline {
    long long t0 = -1;
    __device__ bool expired(long long wait) {
        const long long now = wall_clock64();
        if (t0 < 0) t0 = now;
        return now - t0 > wait;
    }
};

__device__ bool fail_wait(const RedArgs& a) {
    if (fph_lane() == 0) atomicExch(&a.status[0], 5);
    return false;
}

// Copies n published words from src to dst (wavefront-private memory),
// waiting for words not yet written; false after `wait` ticks.
__device__ bool copy_rows_dev(const unsigned* src, int n, unsigned* dst, long long wait) {
    const int lane = fph_lane();
    Deadline dl;
    for (int base = 0; base < n; base += FPH_WAVE) {
        const int i = base + lane;
        bool miss = i < n;
        bool first = true;
        while (true) {
            if (miss) {
                const unsigned v = first ? ld_fast(src + i) : ld_coh(src + i);
                if (v != kUnwritten) {
                    dst[i] = v;
                    miss = false;
                }
            }
            first = false;
            if (!__ballot(miss)) break;
            if (dl.expired(wait)) return false;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    wave_sync();
    return true;
}

// One round trip for columns of up to 63 rows: lane l loads word l of the
// published [len, rows...] at src into v; returns len (waiting for it), or
// -1 after `wait` ticks.  The read may pass the column's end (the pool has
// 64 words of slack past its capacity); those words are ignored.
__device__ int spec_read_dev(const unsigned* src, unsigned& v, long long wait) {
    Deadline dl;
    bool first = true;
    while (true) {
        v = first ? ld_fast(src + fph_lane()) : ld_coh(src + fph_lane());
        first = false;
        const unsigned len = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
        if (len != kUnwritten) return (int)len;
        if (dl.expired(wait)) return -1;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Rows of the column whose first 64 words are in v (spec_read_dev) into
// dst, re-reading words not yet written; false after `wait` ticks.
__device__ bool spec_finish_dev(const unsigned* src, unsigned v, int len, unsigned* dst,
                                long long wait) {
    const int lane = fph_lane();
    Deadline dl;
    bool miss = lane >= 1 && lane <= len;
    while (true) {
        if (miss) {
            if (v != kUnwritten) {
                dst[lane - 1] = v;
                miss = false;
            } else {
                v = ld_coh(src + lane);
            }
        }
        if (!__ballot(miss)) break;
        if (dl.expired(wait)) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    if (len > FPH_WAVE - 1)
        return copy_rows_dev(src + FPH_WAVE, len - (FPH_WAVE - 1), dst + FPH_WAVE - 1, wait);
    wave_sync();
    return true;
}

// Work::half: 0 / 1 = LDS arena halves, 2 / 3 = the wavefront's off-chip
// buffers, -1 = read-only memory (pool or CSR).
//
// w <- w xor (pooled column at ob), whose pivot equals w's.  False on error.
// kDev: the owner column is read with device-coherent loads, waiting for
// words not yet written (device-wide queue), into the stage area or the
// wavefront's third off-chip buffer.
//
// pre (kDev, may be null): lane 0 receives the owner entry of the result's
// pivot, loaded before the merge so that the load overlaps it (the pivot is
// the first element of the symmetric difference, found by a short walk),
// and np its row (0xFFFFFFFF when the result is empty).
template <int H, int O, bool kDev = false>
__device__ bool xor_step(Work& w, const unsigned* ob, unsigned* arena, const RedArgs& a,
                         Alloc& al, int limit = 0x7FFFFFFF, long long wait = 0,
                         unsigned long long* pre = nullptr, unsigned* np = nullptr) {
    unsigned v = 0;
    int nb;
    const unsigned* bs = ob + 1;
    if (kDev) {
        nb = spec_read_dev(ob, v, wait);
        if (nb < 0) return fail_wait(a);
        unsigned* dst = nb <= O ? arena + 2 * H : big_buffer(a, al, 2, nb);
        if (!dst) return false;
        if (!spec_finish_dev(ob, v, nb, dst, wait)) return fail_wait(a);
        bs = dst;
    } else {
        nb = read_line(ob, v);
        if (nb <= O) {
            finish_copy(ob, v, nb, arena + 2 * H);
            bs = arena + 2 * H;
        }
    }
    const unsigned p_old = w.ptr[0];
    const int na = w.len;
    if (kDev && pre) {
        unsigned long long pv = kNoOwner;
        unsigned q = 0xFFFFFFFFu;
        if (fph_lane() == 0) {
            const unsigned* A = w.ptr;
            int i = 1, j = 1;
            while (i < na && j < nb && A[i] == bs[j]) {
                i++;
                j++;
            }
            if (i < na && j < nb) q = min(A[i], bs[j]);
            else if (i < na) q = A[i];
            else if (j < nb) q = bs[j];
            if (q != 0xFFFFFFFFu) pv = load_own(a, q);
        }
        *pre = pv;
        *np = q;
    }
    unsigned* out;
    int out_half;
    const bool on_chip = (w.half == 0 || w.half == 1) && na + nb <= H;
    if (on_chip) {
        out_half = 1 - w.half;
        out = arena + out_half * H;
    } else {
        if (kDev) al.offchip++;
        else if (fph_lane() == 0) atomicAdd(&a.status[4], 1);
        const int bi = w.half == 2 ? 1 : 0;
        out = big_buffer(a, al, bi, na + nb);
        if (!out) return false;
        out_half = 2 + bi;
    }
    int len;
    if (on_chip && nb <= O) len = wave_xor(w.ptr, na, bs, nb, out);
    else len = wave_xor_long(w.ptr, na, bs, nb, out, limit);
    if (!on_chip && len <= H) {  // bring an off-chip result back on chip
        const int h = (w.half == 0 || w.half == 1) ? 1 - w.half : 0;
        for (int i = fph_lane(); i < len; i += FPH_WAVE) arena[h * H + i] = out[i];
        wave_sync();
        out = arena + h * H;
        out_half = h;
    }
    if (kDev) al.merged += (unsigned long long)(na + nb + len);
    else if (fph_lane() == 0) count_merged(a, (long long)na + nb + len);
    w.ptr = out;
    w.half = out_half;
    w.len = len;
    // XOR of two columns with the same pivot removes it, so the pivot
    // strictly increases; anything else means inconsistent inputs.
    if (len > 0 && !(w.ptr[0] > p_old)) {
        if (fph_lane() == 0) atomicExch(&a.status[0], 3);
        return false;
    }
    return true;
}

// Copy the working column into the pool as [len, rows...]; returns the
// offset or -1.
__device__ long long publish(const Work& w, const RedArgs& a, Alloc& al) {
    const long long at = pool_alloc(a, al, w.len + 1);
    if (at < 0) return -1;
    if (fph_lane() == 0) a.pool[at] = (unsigned)w.len;
    wave_copy(w.ptr, w.len, a.pool + at + 1);
    return at;
}

__device__ __forceinline__ void note_len(const RedArgs& a, int len) {
    if (fph_lane() == 0 && len > 0) atomicMax(&a.status[3], len);
}

// Apparent columns are laid out at pool offset off[c] + c as [len, rows...]
// (no allocation), and own[] records them.  Dynamic allocations start after
// this region.
__global__ void k_publish_apparent(RedArgs a, long long K, int set_own) {
    a.lines = lines_of(a);
    const int wpb = blockDim.x / FPH_WAVE;
    const long long nw = (long long)gridDim.x * wpb;
    for (long long c = (long long)blockIdx.x * wpb + threadIdx.x / FPH_WAVE; c < K; c += nw) {
        if (a.kind[c] != 2) continue;
        const int lo = a.off[c], len = a.off[c + 1] - lo;
        const long long at = (long long)lo + c;
        if (fph_lane() == 0) {
            a.pool[at] = (unsigned)len;
            if (set_own)
                a.own[(long long)a.data[lo] - a.row_base] =
                    ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
        }
        for (int i = fph_lane(); i < len; i += FPH_WAVE) a.pool[at + 1 + i] = a.data[lo + i];
        if (a.lines && set_own) {
            const int k = fph_lane();
            const long long row = (long long)a.data[lo] - a.row_base;
            if (k == 0) write_line_word(a, row, 0, (int)c, (unsigned)len);
            else if (k < kLineWords && k <= len) write_line_word(a, row, k, (int)c, a.data[lo + k - 1]);
        }
    }
}

// k_publish_apparent without owners or lines, for many columns: a wavefront
// takes 64 consecutive columns (kinds and offsets in one coalesced round
// trip), writes their headers, and copies their entries -- contiguous in
// `data` -- 64 x kPubUnroll at a time, each entry's column found by a
// binary search over the lanes' offsets.  A wavefront per column waited
// for a kind, then offsets, then entries for every column.
constexpr int kPubUnroll = 4;
__global__ void k_publish_apparent_flat(RedArgs a, long long K) {
    const int lane = fph_lane();
    const long long nw = (long long)gridDim.x * (blockDim.x / FPH_WAVE);
    for (long long c0 = ((long long)blockIdx.x * (blockDim.x / FPH_WAVE) + threadIdx.x / FPH_WAVE) *
                        FPH_WAVE;
         c0 < K; c0 += nw * FPH_WAVE) {
        const long long c = c0 + lane;
        const bool in = c < K;
        const int lo = in ? a.off[c] : 0, hi = in ? a.off[c + 1] : 0;
        const int app = in && a.kind[c] == 2;
        if (app) a.pool[(long long)lo + c] = (unsigned)(hi - lo);
        if (!__ballot(app)) continue;
        const int nin = (int)min<long long>(FPH_WAVE, K - c0);
        const int b0 = __shfl(lo, 0), b1 = __shfl(hi, nin - 1);
        for (int base = b0; base < b1; base += FPH_WAVE * kPubUnroll) {
            unsigned v[kPubUnroll];
            long long dst[kPubUnroll];
#pragma unroll
            for (int u = 0; u < kPubUnroll; u++) {
                const int i = base + u * FPH_WAVE + lane;
                int j = 0;  // the last column of the batch starting at or before i
#pragma unroll
                for (int step = FPH_WAVE / 2; step; step >>= 1) {
                    const int t = j + step;
                    const int lt = __shfl(lo, t < nin ? t : 0);
                    if (t < nin && lt <= i) j = t;
                }
                const int app_j = __shfl(app, j);  // every lane: a cross-lane read
                const bool take = i < b1 && app_j;
                dst[u] = take ? (long long)i + c0 + j + 1 : -1;
                v[u] = take ? a.data[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kPubUnroll; u++)
                if (dst[u] >= 0) a.pool[dst[u]] = v[u];
        }
    }
}

// Phase A: one wavefront per residual column, chaining apparent owners.
__global__ __launch_bounds__(kPhaseAWaves * FPH_WAVE) void k_reduce_chain(
    RedArgs a, const int* __restrict__ res_list, const int* __restrict__ res_count) {
    extern __shared__ unsigned smem[];
    const int wid = threadIdx.x / FPH_WAVE;
    unsigned* arena = smem + wid * kArenaA;
    const int R = *res_count;
    const long long nw = (long long)gridDim.x * kPhaseAWaves;
    int nxor = 0;
    Alloc al = make_alloc(128);
    for (long long r = (long long)blockIdx.x * kPhaseAWaves + wid; r < R; r += nw) {
        if (a.status[0]) return;
        const int c = res_list[r];
        const int lo = a.off[c], n0 = a.off[c + 1] - lo;
        Work w;
        w.len = n0;
        if (n0 <= kHalfA) {
            wave_copy(a.data + lo, n0, arena);
            w.ptr = arena;
            w.half = 0;
        } else {
            w.ptr = a.data + lo;  // read-only; XOR results go elsewhere
            w.half = -1;
        }
        while (w.len > 0) {
            const unsigned long long v = a.own[(long long)w.ptr[0] - a.row_base];
            if (v == kNoOwner) break;
            if (!xor_step<kHalfA, kStageA>(w, a.pool + (v & 0xFFFFFFFFull), arena, a, al))
                return;
            nxor++;
        }
        note_len(a, w.len);
        const long long at = publish(w, a, al);
        if (at < 0) return;
        if (fph_lane() == 0) {
            a.desc[c] = (unsigned)at;
            a.pivot[c] = w.len > 0 ? (int)w.ptr[0] : -1;
        }
    }
    if (fph_lane() == 0 && nxor) atomicAdd(&a.status[2], nxor);
}

// Phase B work item: reduce column c against current owners, then claim.
// Returns the displaced column (>= 0), -1 if none, -2 on error.
// kAgent: claimers run in several workgroups (on different CUs), so the
// publish -> claim and claim -> read orders need device (agent) scope, and a
// reader must acquire before reading an owner's pooled content.
template <bool kAgent>
__device__ int reduce_claim(const RedArgs& a, int c, unsigned* arena, int* nxor, Alloc& al) {
    Work w;
    work_load(w, a.pool + a.desc[c], arena, kHalfB);
    while (true) {
        if (w.len == 0) {
            if (fph_lane() == 0) a.pivot[c] = -1;
            return -1;
        }
        const unsigned p = w.ptr[0];
        unsigned long long v = load_own(a, p);
        if (kAgent && v != kNoOwner) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if ((int)(v >> 32) - 1 <= c) {
            const long long at = publish(w, a, al);
            if (at < 0) return -2;
            // desc / pivot before the claim: once the claim lands, a higher
            // column may displace c at once and another wavefront may take c
            // up (queue mode); a lost claim rewrites them later
            if (fph_lane() == 0) {
                a.desc[c] = (unsigned)at;
                a.pivot[c] = (int)p;
            }
            if (kAgent) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const unsigned long long mine =
                ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
            unsigned long long old = 0;
            if (fph_lane() == 0) old = atomicMax(&a.own[(long long)p - a.row_base], mine);
            old = __shfl(old, 0);
            if (kAgent) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int oc = (int)(old >> 32) - 1;
            if (oc < c) {
                note_len(a, w.len);
                return oc;  // displaced owner (or -1)
            }
            v = old;  // a higher-rank column claimed p first: reduce by it
        }
        if (!xor_step<kHalfB, kStageB>(w, a.pool + (v & 0xFFFFFFFFull), arena, a, al))
            return -2;
        (*nxor)++;
    }
}

// Phase B: one workgroup per segment of columns (a point cloud of a batch;
// the whole complex otherwise).  Segments never share rows, so they reduce
// independently.  The 16 wavefronts take columns from a queue: first the
// segment's residual columns in descending rank, then the owners displaced
// by claims, in the order of displacement (a ring of (sequence << 32 |
// column) words in the segment's column range: at most as many columns are
// pending as the segment has, so a slot is never reused while unread).  A
// wavefront with nothing to take waits for its slot and leaves when no
// column is pending; the waits are bounded (status 5).  All communication is
// inside one workgroup, so workgroup-scope ordering suffices.
__global__ __launch_bounds__(kPhaseBWaves * FPH_WAVE) void k_reduce_queue(
    RedArgs a, const unsigned* __restrict__ order, const int* __restrict__ res_count,
    const int* __restrict__ seg_cols, unsigned long long* __restrict__ ring,
    long long spin_limit) {
    extern __shared__ unsigned smem[];
    __shared__ int s_lo, s_cnt, s_abort;
    __shared__ unsigned s_head, s_dhead, s_dtail;
    __shared__ int s_pending;
    const int tid = threadIdx.x;
    const int wid = tid / FPH_WAVE;
    const int lane = fph_lane();
    unsigned* arena = smem + wid * kArenaB;
    const int c0 = seg_cols ? seg_cols[blockIdx.x] : 0;
    const int c1 = seg_cols ? seg_cols[blockIdx.x + 1] : (int)a.K;
    if (tid == 0) {
        // order[0..R) holds the residual columns in descending order; this
        // segment's columns [c0, c1) form the run order[lo..hi).
        const int R = *res_count;
        int lo = 0, hi = R;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)order[mid] >= c1) lo = mid + 1;
            else hi = mid;
        }
        int lo2 = lo, hi2 = R;
        while (lo2 < hi2) {
            const int mid = (lo2 + hi2) >> 1;
            if ((int)order[mid] >= c0) lo2 = mid + 1;
            else hi2 = mid;
        }
        s_lo = lo;
        s_cnt = lo2 - lo;
        s_abort = a.status[0];
        s_head = 0;
        s_dhead = 0;
        s_dtail = 0;
        s_pending = lo2 - lo;
    }
    __syncthreads();
    if (s_abort || s_cnt == 0) return;
    const unsigned cap = (unsigned)(c1 - c0);  // >= pending columns
    unsigned long long* q = ring + c0;
    int nxor = 0;
    Alloc al = make_alloc(4096);
    while (true) {
        if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
        unsigned k = 0;
        if (lane == 0) k = atomicAdd(&s_head, 1u);
        k = __shfl(k, 0);
        int c = -1;
        if (k < (unsigned)s_cnt) {
            c = (int)order[s_lo + k];
        } else {
            unsigned slot = 0;
            if (lane == 0) slot = atomicAdd(&s_dhead, 1u);
            slot = __shfl(slot, 0);
            // bounded wait: give up (status 5) after spin_limit wall-clock
            // ticks without progress (no change of the pending count or
            // the ring tail) -- other wavefronts may legitimately work on a
            // long chain meanwhile
            // (the clock is read every 16th round; see Deadline)
            unsigned long long t_last = 0;
            bool timing = false;
            int last_pend = -1;
            unsigned last_tail = 0xFFFFFFFFu;
            for (unsigned it = 0;; it++) {
                unsigned long long v = 0;
                int pend = 0, ab = 0;
                unsigned tail = 0;
                if (lane == 0) {
                    v = __hip_atomic_load(&q[slot % cap], __ATOMIC_ACQUIRE,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
                    pend = __hip_atomic_load(&s_pending, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                    ab = __hip_atomic_load(&s_abort, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    tail = __hip_atomic_load(&s_dtail, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                v = __shfl(v, 0);
                pend = __shfl(pend, 0);
                ab = __shfl(ab, 0);
                tail = __shfl(tail, 0);
                if ((unsigned)(v >> 32) == slot) {
                    c = (int)(v & 0xFFFFFFFFull);
                    break;
                }
                if (pend == 0 || ab) break;
                if ((it & 15) == 15) {
                    const unsigned long long now = wall_clock64();
                    if (!timing || pend != last_pend || tail != last_tail) {
                        timing = true;
                        last_pend = pend;
                        last_tail = tail;
                        t_last = now;
                    }
                    if (now - t_last > (unsigned long long)spin_limit) {
                        if (lane == 0) {
                            atomicExch(&a.status[0], 5);
                            atomicExch(&s_abort, 1);
                        }
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (c < 0) break;
        }
        int d = -1;
        if (a.pivot[c] >= 0) {  // zero after phase A otherwise
            d = reduce_claim<false>(a, c, arena, &nxor, al);
            if (d == -2) {
                if (lane == 0) {
                    atomicExch(&s_abort, 1);
                    atomicSub(&s_pending, 1);
                }
                break;
            }
        }
        if (lane == 0) {
            if (d >= 0) {
                atomicAdd(&s_pending, 1);
                const unsigned t = atomicAdd(&s_dtail, 1u);
                __hip_atomic_store(&q[t % cap], ((unsigned long long)t << 32) | (unsigned)d,
                                   __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            atomicSub(&s_pending, 1);
        }
    }
    if (lane == 0 && nxor) atomicAdd(&a.status[2], nxor);
    if (tid == 0) atomicMax(&a.status[1], 1);
}

// ---- device-wide queue ----------------------------------------------------
// The queue of k_reduce_queue spread over several workgroups (one segment
// only).  Nothing is ordered by fences: pool words are write-once over
// kUnwritten and read with device-coherent waits (above); own[] changes only
// by atomicMax and is read coherently; desc[c] may be read stale, which only
// restarts c from an older published state of c (every state of c is c's
// column plus owners of higher rank, so any of them reduces to a valid
// column).  pivot[] is not written here: k_final_* rebuild pivot[] and
// desc[] from the final own[] afterwards.

__device__ long long publish_dev(const Work& w, const RedArgs& a, Alloc& al) {
    const long long at = pool_alloc(a, al, w.len + 1);
    if (at < 0) return -1;
    for (int i = fph_lane(); i < w.len; i += FPH_WAVE) st_coh(a.pool + at + 1 + i, w.ptr[i]);
    if (fph_lane() == 0) st_coh(a.pool + at, (unsigned)w.len);
    wave_sync();
    return at;
}

// Reduce column c, whose current state is published at pool offset src,
// and claim its pivot.  Returns the displaced column (>= 0), -1 if none, -2
// on error.
__device__ int reduce_claim_dev(const RedArgs& a, int c, unsigned src, unsigned* arena,
                                int* nxor, Alloc& al, long long wait) {
    Work w;
    unsigned v0;
    const int len0 = spec_read_dev(a.pool + src, v0, wait);
    if (len0 < 0) {
        fail_wait(a);
        return -2;
    }
    if (len0 <= kHalfB) {
        w.ptr = arena;
        w.half = 0;
    } else {
        unsigned* b = big_buffer(a, al, 0, len0);
        if (!b) return -2;
        w.ptr = b;
        w.half = 2;
    }
    if (!spec_finish_dev(a.pool + src, v0, len0, const_cast<unsigned*>(w.ptr), wait)) {
        fail_wait(a);
        return -2;
    }
    w.len = len0;
    unsigned long long pre = kNoOwner;
    unsigned np = 0xFFFFFFFFu;
    bool have_pre = false;
    while (true) {
        if (w.len == 0) return -1;
        const unsigned p = w.ptr[0];
        unsigned long long v;
        // the owner entry prefetched during the last merge, if it is p's
        if (have_pre && __shfl(np, 0) == p) v = __shfl(pre, 0);
        else v = load_own(a, p);
        have_pre = false;
        if ((int)(v >> 32) - 1 <= c) {
            const long long at = publish_dev(w, a, al);
            if (at < 0) return -2;
            if (fph_lane() == 0) st_coh(&a.desc[c], (unsigned)at);
            const unsigned long long mine =
                ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
            unsigned long long old = 0;
            if (fph_lane() == 0) old = atomicMax(&a.own[(long long)p - a.row_base], mine);
            old = __shfl(old, 0);
            const int oc = (int)(old >> 32) - 1;
            if (oc < c) {
                al.maxlen = max(al.maxlen, w.len);
                if (oc >= 0 && fph_lane() == 0)
                    st_coh(&a.desc[oc], (unsigned)(old & 0xFFFFFFFFull));
                return oc;
            }
            v = old;
        }
        if (!xor_step<kHalfB, kStageB, true>(w, a.pool + (v & 0xFFFFFFFFull), arena, a, al,
                                             0x7FFFFFFF, wait, &pre, &np))
            return -2;
        have_pre = true;
        (*nxor)++;
    }
}

// ---- bitmap working column (device-wide queue, rows fit the arena) -------
// When the rows [row_base, row_base + nrows) fit the wavefront's arena as a
// bitmap, the working column lives there: an addition flips the owner's rows
// (LDS atomic XOR, one row per lane) and the pivot moves to the next set bit
// at or above the old one (pivots only increase).  The cost of an addition
// is then the owner's length, not the working column's (long working
// columns of a few hundred rows meet owners of a few rows).

struct BmWork {
    unsigned* bm;   // wavefront arena
    int lo_w;       // first word that may hold a set bit
    int hi_w;       // last word that may hold a set bit
    int len;        // set bits
    unsigned piv;   // lowest set row (global rank) when len > 0
};

// Owner entry read through the L2 (may be stale; see reduce_claim_dev_bm).
__device__ __forceinline__ unsigned long long load_own_fast(const RedArgs& a, unsigned p) {
    return __hip_atomic_load(&a.own[(long long)p - a.row_base], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until every lane with `need` holds a written word in x (first read
// already issued by the caller); false after `wait` ticks.
__device__ __forceinline__ bool settle_word(const unsigned* p, bool need, unsigned& x,
                                            Deadline& dl, long long wait) {
    bool miss = need && x == kUnwritten;
    while (__ballot(miss)) {
        if (dl.expired(wait)) return false;
        __builtin_amdgcn_s_sleep(1);
        if (miss) {
            x = ld_coh(p);
            miss = x == kUnwritten;
        }
    }
    return true;
}

// Flips the rows of the published column at src ([n, rows...]; its first 64
// words already in v, lane l holding word l) into the bitmap.  Returns the
// lowest row flipped (all lanes), or 0xFFFFFFFF after `wait` ticks.
__device__ unsigned bm_flip_published(const RedArgs& a, BmWork& w, const unsigned* src,
                                      unsigned v, int n, long long wait) {
    const int lane = fph_lane();
    Deadline dl;
    int delta = 0, hi = -1;  // wave-uniform
    unsigned first = 0;
    for (int base = 0; base <= n; base += FPH_WAVE) {
        const int i = base + lane;            // word index in [len, rows...]
        const bool need = i >= 1 && i <= n;
        unsigned x = base == 0 ? v : (need ? ld_fast(src + i) : 0u);
        if (!settle_word(src + i, need, x, dl, wait)) return 0xFFFFFFFFu;
        if (base == 0) first = (unsigned)__builtin_amdgcn_readlane((int)x, 1);
        const unsigned r = x - (unsigned)a.row_base;
        const int wd = (int)(r >> 5);
        bool set = false;
        if (need) {
            const unsigned bit = 1u << (r & 31);
            set = (atomicXor(&w.bm[wd], bit) & bit) == 0u;
        }
        // rows ascend: the chunk's last row holds its highest word
        const unsigned long long mn = __ballot(need);
        delta += 2 * __popcll(__ballot(set)) - __popcll(mn);
        if (mn) hi = max(hi, __builtin_amdgcn_readlane(wd, 63 - __clzll((long long)mn)));
    }
    w.len += delta;
    w.hi_w = max(w.hi_w, hi);
    wave_sync();
    return first;
}

// Flips n rows of an immutable list (the CSR columns) into the bitmap;
// returns the lowest row.
__device__ unsigned bm_flip_list(const RedArgs& a, BmWork& w, const unsigned* rows, int n) {
    int delta = 0, hi = -1;  // wave-uniform
    for (int base = 0; base < n; base += FPH_WAVE) {
        const int i = base + fph_lane();
        const bool need = i < n;
        int wd = 0;
        bool set = false;
        if (need) {
            const unsigned r = rows[i] - (unsigned)a.row_base;
            wd = (int)(r >> 5);
            const unsigned bit = 1u << (r & 31);
            set = (atomicXor(&w.bm[wd], bit) & bit) == 0u;
        }
        // rows ascend: the chunk's last row holds its highest word
        const unsigned long long mn = __ballot(need);
        delta += 2 * __popcll(__ballot(set)) - __popcll(mn);
        hi = max(hi, __builtin_amdgcn_readlane(wd, 63 - __clzll((long long)mn)));
    }
    w.len += delta;
    w.hi_w = max(w.hi_w, hi);
    wave_sync();
    return rows[0];
}

// Lowest set bit at or above the current pivot (len > 0).
__device__ void bm_next_pivot(const RedArgs& a, BmWork& w) {
    const int lane = fph_lane();
    for (int wd = (int)((w.piv - (unsigned)a.row_base) >> 5); wd <= w.hi_w; wd += FPH_WAVE) {
        const int i = wd + lane;
        const unsigned x = i <= w.hi_w ? w.bm[i] : 0u;
        const unsigned long long b = __ballot(x != 0u);
        if (b) {
            const int l = __ffsll((long long)b) - 1;
            const unsigned xw = (unsigned)__builtin_amdgcn_readlane((int)x, l);
            w.piv = (unsigned)a.row_base + 32u * (unsigned)(wd + l) + (unsigned)(__ffs(xw) - 1);
            return;
        }
    }
}

// Adds a published column of n <= 63 rows whose words [len, rows...] are
// all in v (lane l holds word l, every word written) and moves the pivot to
// the lowest set bit, in one LDS round trip: the flips and the first window
// of the pivot scan are issued back to back, and a wavefront's LDS
// operations complete in issue order, so the window reads see the flips.
// The added column's pivot is the current one, so every row it flips is at
// or above it and the lowest set bit lies at or above the current pivot's
// word.  Returns 1, 0 when the working column became zero, or -1 when no
// set bit lies in [pivot word, hi_w] although len > 0 (inconsistent).
__device__ int bm_add_short(const RedArgs& a, BmWork& w, unsigned v, int n) {
    const int lane = fph_lane();
    const bool need = lane >= 1 && lane <= n;
    const unsigned r = v - (unsigned)a.row_base;
    const int wd = need ? (int)(r >> 5) : 0;
    const unsigned bit = 1u << (r & 31);
    // rows ascend: lane n holds the highest word
    w.hi_w = max(w.hi_w, __builtin_amdgcn_readlane(wd, n));
    int base = (int)((w.piv - (unsigned)a.row_base) >> 5);
    unsigned old = 0u;
    if (need) old = atomicXor(&w.bm[wd], bit);
    asm volatile("" ::: "memory");  // keep the window reads after the flips
    unsigned x = base + lane <= w.hi_w ? w.bm[base + lane] : 0u;
    w.len += 2 * __popcll(__ballot(need && (old & bit) == 0u)) - n;
    if (w.len == 0) return 0;
    unsigned long long b = __ballot(x != 0u);
    while (!b) {  // len > 0: a set bit lies at or below hi_w
        base += FPH_WAVE;
        if (base > w.hi_w) return -1;
        x = base + lane <= w.hi_w ? w.bm[base + lane] : 0u;
        b = __ballot(x != 0u);
    }
    const int l = __ffsll((long long)b) - 1;
    const unsigned xw = (unsigned)__builtin_amdgcn_readlane((int)x, l);
    w.piv = (unsigned)a.row_base + 32u * (unsigned)(base + l) + (unsigned)(__ffs(xw) - 1);
    return 1;
}

// Publishes the set rows as [len, rows...] (write-once words) and, with
// owner lines, the line of the pivot row tagged with column c; -1 on error.
__device__ long long bm_publish(const RedArgs& a, const BmWork& w, Alloc& al, int c) {
    const int lane = fph_lane();
    const long long at = pool_alloc(a, al, w.len + 1);
    if (at < 0) return -1;
    unsigned* dst = a.pool + at + 1;
    int done = 0;
    for (int wd = (int)((w.piv - (unsigned)a.row_base) >> 5); wd <= w.hi_w && done < w.len;
         wd += FPH_WAVE) {
        const int i = wd + lane;
        unsigned x = i <= w.hi_w ? w.bm[i] : 0u;
        const int cnt = __popc(x);
        int incl = cnt;
        for (int o = 1; o < FPH_WAVE; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        int pos = done + incl - cnt;
        while (x && pos < w.len) {  // pos < len always; the bound keeps writes inside
            const int b = __ffs(x) - 1;
            const unsigned row = (unsigned)a.row_base + 32u * (unsigned)i + (unsigned)b;
            st_coh(dst + pos, row);
            if (a.lines && pos < kLineWords - 1)
                write_line_word(a, (long long)w.piv - a.row_base, pos + 1, c, row);
            pos++;
            x &= x - 1u;
        }
        done += __shfl(incl, FPH_WAVE - 1);
    }
    if (lane == 0) {
        st_coh(a.pool + at, (unsigned)w.len);
        if (a.lines) write_line_word(a, (long long)w.piv - a.row_base, 0, c, (unsigned)w.len);
    }
    wave_sync();
    return at;
}

__device__ void bm_clear(BmWork& w) {
    for (int i = w.lo_w + fph_lane(); i <= w.hi_w; i += FPH_WAVE) w.bm[i] = 0u;
    wave_sync();
}

// reduce_claim_dev with the bitmap working column (bm: the wavefront's
// arena, all zero on entry and on return).  Owner entries are read through
// the L2 and may be stale: a stale owner is still a column of higher rank
// with the same pivot (a valid addition), a stale "no owner" leads to a
// claim whose atomicMax returns the true owner, and every addition raises
// the pivot, so the loop ends.
// csr: c starts from its cofacet column in the CSR (not yet published).
__device__ int reduce_claim_dev_bm(const RedArgs& a, int c, unsigned src, bool csr, unsigned* bm,
                                   int* nxor, Alloc& al, long long wait) {
    const int lane = fph_lane();
    BmWork w;
    w.bm = bm;
    w.len = 0;
    w.hi_w = -1;
    if (csr) {
        const int lo = a.off[c], n0 = a.off[c + 1] - lo;
        if (n0 == 0) return -1;
        w.piv = bm_flip_list(a, w, a.data + lo, n0);
    } else {
        unsigned v0;
        const int n0 = spec_read_dev(a.pool + src, v0, wait);
        if (n0 < 0) {
            fail_wait(a);
            return -2;
        }
        if (n0 == 0) return -1;
        w.piv = bm_flip_published(a, w, a.pool + src, v0, n0, wait);
        if (w.piv == 0xFFFFFFFFu) {
            fail_wait(a);
            return -2;
        }
    }
    w.lo_w = (int)((w.piv - (unsigned)a.row_base) >> 5);
    while (true) {
        if (w.len == 0) {
            bm_clear(w);
            return -1;
        }
        const unsigned p = w.piv;
        // the owner entry and, in the same round trip, the owner line
        unsigned long long lw = 0;
        if (a.lines && lane < kLineWords)
            lw = __hip_atomic_load(&a.lines[((long long)p - a.row_base) * kLineWords + lane],
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        unsigned long long v = load_own_fast(a, p);
        if ((int)(v >> 32) - 1 <= c) {
            const long long at = bm_publish(a, w, al, c);
            if (at < 0) return -2;
            if (lane == 0) st_coh(&a.desc[c], (unsigned)at);
            const unsigned long long mine =
                ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
            unsigned long long old = 0;
            if (lane == 0) old = atomicMax(&a.own[(long long)p - a.row_base], mine);
            old = __shfl(old, 0);
            const int oc = (int)(old >> 32) - 1;
            if (oc < c) {
                al.maxlen = max(al.maxlen, w.len);
                if (oc >= 0 && lane == 0) st_coh(&a.desc[oc], (unsigned)(old & 0xFFFFFFFFull));
                bm_clear(w);
                return oc;
            }
            v = old;
        }
        const unsigned* ob = a.pool + (v & 0xFFFFFFFFull);
        // The line serves when every word up to the length carries the tag
        // of the owner just read: (owner, pivot) names one published state
        // of the owner (its pivot only rises), so the line holds that
        // state's first rows.  Otherwise (stale, torn, longer than the line)
        // the column itself is read.
        unsigned vb = 0;
        int nb = -1;
        bool complete = false;  // every word of [len, rows...] in vb
        if (a.lines && (unsigned long long)v != kNoOwner) {
            const unsigned tag = (unsigned)(v >> 32);
            const unsigned lo32 = (unsigned)lw, hi32 = (unsigned)(lw >> 32);
            const unsigned t0 = (unsigned)__builtin_amdgcn_readlane((int)hi32, 0);
            const unsigned l0 = (unsigned)__builtin_amdgcn_readlane((int)lo32, 0);
            if (t0 == tag && l0 <= (unsigned)(kLineWords - 1)) {
                const bool need = lane >= 1 && lane <= (int)l0;
                if (!__ballot(need && hi32 != tag)) {
                    nb = (int)l0;
                    vb = lo32;
                    complete = true;
                }
            }
        }
        if (nb < 0) {
            nb = spec_read_dev(ob, vb, wait);
            if (nb < 0) {
                fail_wait(a);
                return -2;
            }
        }
        const int na = w.len;
        if (nb <= FPH_WAVE - 1) {
            if (!complete) {  // rows read from the pool: wait for unwritten words
                Deadline dl;
                if (!settle_word(ob + lane, lane >= 1 && lane <= nb, vb, dl, wait)) {
                    fail_wait(a);
                    return -2;
                }
            }
            if (bm_add_short(a, w, vb, nb) < 0) {
                if (lane == 0) atomicExch(&a.status[0], 3);
                return -2;
            }
        } else {
            if (bm_flip_published(a, w, ob, vb, nb, wait) == 0xFFFFFFFFu) {
                fail_wait(a);
                return -2;
            }
            if (w.len > 0) bm_next_pivot(a, w);
        }
        // the owner's pivot is p, so p is cleared and the pivot must rise
        if (w.len > 0 && !(w.piv > p)) {
            if (lane == 0) atomicExch(&a.status[0], 3);
            return -2;
        }
        al.merged += (unsigned long long)(na + nb + w.len);
        (*nxor)++;
    }
}

// qs: [0] next residual index, [1] next ring slot to read, [2] ring tail,
// [3] pending columns, [4] abort, [5] XCD taking part (-1: none yet).
// BITMAP: working columns as bitmaps in the wavefront's arena (rows fit
// it), else sorted lists; two instantiations, so that the bitmap kernel
// carries none of the list merges' code and registers.
template <bool BITMAP>
__global__ __launch_bounds__(kPhaseBWaves * FPH_WAVE) void k_reduce_queue_dev(
    RedArgs a, const unsigned* __restrict__ order, const int* __restrict__ res_count,
    unsigned long long* __restrict__ ring, int* __restrict__ qs, long long spin_limit,
    int from_csr) {
    extern __shared__ unsigned smem[];
    __shared__ int s_part;
    a.lines = lines_of(a);
    // Only the workgroups on one XCD take part (the XCD of the first one to
    // start): they share that XCD's L2, where the columns they publish stay,
    // so workgroup-scope reads (ld_fast) mostly hit.  Correctness does not
    // depend on it (write-once words, coherent re-reads).
    if (threadIdx.x == 0) {
        const int xcc = (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 0xF);
        const int lead = atomicCAS(qs + 5, -1, xcc);
        s_part = lead == -1 || lead == xcc;
    }
    __syncthreads();
    if (!s_part) return;
    const int wid = threadIdx.x / FPH_WAVE;
    const int lane = fph_lane();
    unsigned* arena = smem + wid * kArenaB;
    const int R = *res_count;
    const unsigned cap = (unsigned)a.K;  // >= pending columns
    int nxor = 0;
    Alloc al = make_alloc(512);  // many wavefronts: small private chunks
    if (BITMAP) {  // the bitmap starts (and is left) all zero
        for (int i = lane; i < kArenaB; i += FPH_WAVE) arena[i] = 0u;
        wave_sync();
    }
    while (true) {
        if (ld_coh(reinterpret_cast<const unsigned*>(qs + 4))) break;
        unsigned k = 0;
        if (lane == 0) k = atomicAdd(reinterpret_cast<unsigned*>(qs), 1u);
        k = __shfl(k, 0);
        int c = -1;
        if (k < (unsigned)R) {
            c = (int)order[k];
        } else {
            unsigned slot = 0;
            if (lane == 0) slot = atomicAdd(reinterpret_cast<unsigned*>(qs + 1), 1u);
            slot = __shfl(slot, 0);
            // poll the own slot every round and the shared counters (one
            // cache line that every waiting wavefront reads) every 8th, so
            // that idle wavefronts leave the memory path to the working ones
            // the clock is read with the counters, every 64th round
            long long t_last = -1;
            int last_pend = -1;
            unsigned last_tail = 0xFFFFFFFFu;
            for (unsigned it = 0;; it++) {
                unsigned long long v = 0;
                if (lane == 0)
                    v = __hip_atomic_load(&ring[slot % cap], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                v = __shfl(v, 0);
                if ((unsigned)(v >> 32) == slot) {
                    c = (int)(v & 0xFFFFFFFFull);
                    break;
                }
                if ((it & 7) == 7) {
                    int pend = 0, ab = 0;
                    unsigned tail = 0;
                    if (lane == 0) {
                        pend = (int)ld_coh(reinterpret_cast<const unsigned*>(qs + 3));
                        ab = (int)ld_coh(reinterpret_cast<const unsigned*>(qs + 4));
                        tail = ld_coh(reinterpret_cast<const unsigned*>(qs + 2));
                    }
                    pend = __shfl(pend, 0);
                    ab = __shfl(ab, 0);
                    tail = __shfl(tail, 0);
                    if (pend == 0 || ab) break;
                    if ((it & 63) != 63) {
                        __builtin_amdgcn_s_sleep(2);
                        continue;
                    }
                    const long long now = wall_clock64();
                    if (t_last < 0 || pend != last_pend || tail != last_tail) {
                        last_pend = pend;
                        last_tail = tail;
                        t_last = now;
                    }
                    if (now - t_last > spin_limit) {
                        if (lane == 0) {
                            atomicExch(&a.status[0], 5);
                            atomicExch(qs + 4, 1);
                        }
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(2);
            }
            if (c < 0) break;
        }
        int d = -1;
        // from_csr: no phase A ran; desc[c] reads kUnwritten until c first
        // publishes, and then c starts from its CSR column (never zero).
        // Otherwise pivot < 0 marks columns phase A reduced to zero.
        if (from_csr || a.pivot[c] >= 0) {
            const unsigned src = ld_coh(&a.desc[c]);
            const bool csr = from_csr && src == kUnwritten;
            if constexpr (BITMAP)
                d = reduce_claim_dev_bm(a, c, src, csr, arena, &nxor, al, spin_limit);
            else
                d = reduce_claim_dev(a, c, src, arena, &nxor, al, spin_limit);
            if (d == -2) {
                if (lane == 0) {
                    atomicExch(qs + 4, 1);
                    atomicSub(qs + 3, 1);
                }
                break;
            }
        }
        if (lane == 0) {
            if (d >= 0) {
                atomicAdd(qs + 3, 1);
                const unsigned t = atomicAdd(reinterpret_cast<unsigned*>(qs + 2), 1u);
                __hip_atomic_store(&ring[t % cap], ((unsigned long long)t << 32) | (unsigned)d,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            atomicSub(qs + 3, 1);
        }
    }
    if (lane == 0) {
        if (nxor) atomicAdd(&a.status[2], nxor);
        if (al.merged) count_merged(a, (long long)al.merged);
        if (al.offchip) atomicAdd(&a.status[4], al.offchip);
        if (al.maxlen) atomicMax(&a.status[3], al.maxlen);
    }
    if (threadIdx.x == 0) atomicMax(&a.status[1], 1);
}

// Pool space not handed out yet reads kUnwritten: [*pool_ptr, end).
__global__ void k_fill_unwritten(unsigned* __restrict__ pool,
                                 const unsigned long long* __restrict__ pool_ptr,
                                 long long end) {
    const long long start = (long long)*pool_ptr;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = start + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < end;
         i += stride)
        pool[i] = kUnwritten;
}

__global__ void k_queue_state_init(int* __restrict__ qs, const int* __restrict__ res_count,
                                   const int* __restrict__ status) {
    if (threadIdx.x == 0) {
        qs[0] = 0;
        qs[1] = 0;
        qs[2] = 0;
        qs[3] = *res_count;
        qs[4] = status[0] != 0;
        qs[5] = -1;
    }
}

// Set-up of the device-wide queue with bitmap working columns, one launch:
// the status and pool pointer (k_reduce_init's work), owner lines zero, the
// ring empty, desc[] unpublished, pivot[] reset (no phase A ran and the
// queue never reads pivot[]; k_final_own rebuilds it), the unallocated pool
// reading kUnwritten, and the queue state.
__global__ void k_queue_prep(RedArgs a, unsigned long long* __restrict__ ring,
                             int* __restrict__ qs, const int* __restrict__ res_count,
                             long long fill_end) {
    unsigned long long* lines = lines_of(a);
    const long long nl = lines ? a.nrows * kLineWords : 0;
    // the pool's first free word, as k_reduce_init sets it: apparent region
    // (off[K] + K words), one line of padding, then the owner lines
    unsigned long long pp =
        ((unsigned long long)a.off[a.K] + a.K + kLine - 1) / kLine * kLine + kLine;
    if (lines) pp += 2ull * kLineWords * (unsigned long long)a.nrows;
    const long long start = (long long)pp;
    const long long nf = fill_end > start ? fill_end - start : 0;
    const long long n = max(max(nl, a.K), nf);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (i < nl) lines[i] = 0ull;
        if (i < a.K) {
            ring[i] = ~0ull;
            a.desc[i] = kUnwritten;
            if (a.kind[i] == 2 || a.kind[i] == 3) a.pivot[i] = -1;
        }
        if (i < nf) a.pool[start + i] = kUnwritten;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int i = 0; i < 8; i++) a.status[i] = 0;  // k_reduce_init's reset
        *a.pool_ptr = pp;
        qs[0] = 0;
        qs[1] = 0;
        qs[2] = 0;
        qs[3] = *res_count;
        qs[4] = 0;
        qs[5] = -1;
    }
}

// After the device-wide queue: every nonzero column owns exactly its pivot
// row, so own[] alone gives pivot[] and desc[] (reset first, then set).
__global__ void k_final_reset(RedArgs a) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.K && (a.kind[c] == 2 || a.kind[c] == 3)) a.pivot[c] = -1;
}

__global__ void k_final_own(RedArgs a) {
    long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) {  // k_reduce_fini's work: the pool words used, in lines
        const unsigned long long lines = a.pool_ptr[0] / kLine;
        a.status[5] = lines > 0x7FFFFFFFull ? 0x7FFFFFFF : (int)lines;
    }
    if (r >= a.nrows) return;
    const unsigned long long v = a.own[r];
    if (v == kNoOwner) return;
    const long long c = (long long)(v >> 32) - 1;
    a.pivot[c] = (int)(r + a.row_base);
    a.desc[c] = (unsigned)(v & 0xFFFFFFFFull);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, long long n, unsigned long long v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------
// Replay of the reference's reduction from the final owners.
//
// _general_cohomology_reduce_v2 (cohomology_general.py:255-366) takes the
// residual columns in descending order and adds to each the stored column
// of its current pivot's owner until the pivot has none; a column is stored
// (and its pivot gets an owner) only while the store has room, and working
// columns stop at MAX_COL rows (_numba_utils.py:44-63).  While every column
// fits, that is the standard reduction, whose pairs are those of any
// complete reduction: the owner a column meets at pivot p is the column
// whose final pivot is p, as fph_reduce leaves it in own[].  So each of the
// reference's columns can be rebuilt on its own wavefront -- from its CSR
// column, adding the final owner's column (the CSR column of an apparent
// owner, else the column this replay published for it: owners come earlier
// in the order, so their wavefronts are already running) until the pivot
// is the column's own final pivot.  The replay reports what the reference's
// store would hold and the longest working column: when both stay within
// its limits, fph_reduce's pairs are the reference's.

// Waits until desc[k] is published; kUnwritten on abort (qs[4]) or timeout
// (status 5).
__device__ unsigned replay_wait_desc(const RedArgs& a, int k, const int* qs, long long wait) {
    unsigned src = __shfl(ld_fast(&a.desc[k]), 0);
    Deadline dl;
    while (src == kUnwritten) {
        if (__shfl(ld_coh(reinterpret_cast<const unsigned*>(qs + 4)), 0)) return kUnwritten;
        if (dl.expired(wait)) {
            fail_wait(a);
            return kUnwritten;
        }
        __builtin_amdgcn_s_sleep(2);
        src = __shfl(ld_coh(&a.desc[k]), 0);
    }
    return src;
}

// Rebuilds the reference's reduced column c in the bitmap bm (zero on entry
// and on return) and publishes it when non-zero.  Returns its length, or -1
// (status set, or an abort seen).
__device__ int replay_column(const RedArgs& a, int c, unsigned* bm, Alloc& al, const int* qs,
                             int& adds, int& longest, long long wait) {
    const int lo = a.off[c], n0 = a.off[c + 1] - lo;
    if (n0 <= 0) return 0;  // no cofacets: essential, not stored
    BmWork w;
    w.bm = bm;
    w.len = 0;
    w.hi_w = -1;
    w.piv = bm_flip_list(a, w, a.data + lo, n0);
    w.lo_w = (int)((w.piv - (unsigned)a.row_base) >> 5);
    longest = max(longest, n0);
    const int target = a.pivot[c];
    while (w.len > 0 && (int)w.piv != target) {
        const unsigned p = w.piv;
        const unsigned long long v = a.own[(long long)p - a.row_base];
        const int k = (int)(v >> 32) - 1;
        const int kk = (v == kNoOwner || k <= c || k >= a.K) ? 0 : (int)a.kind[k];
        if (kk == 2) {  // apparent owner: its CSR column
            const int lk = a.off[k];
            bm_flip_list(a, w, a.data + lk, a.off[k + 1] - lk);
        } else if (kk == 3) {
            const unsigned src = replay_wait_desc(a, k, qs, wait);
            if (src == kUnwritten) return -1;
            unsigned vb;
            const int nb = spec_read_dev(a.pool + src, vb, wait);
            if (nb < 0) {
                fail_wait(a);
                return -1;
            }
            if (bm_flip_published(a, w, a.pool + src, vb, nb, wait) == 0xFFFFFFFFu) {
                fail_wait(a);
                return -1;
            }
        } else {  // no earlier owner of a pivot the column passes: inconsistent
            if (fph_lane() == 0) atomicExch(&a.status[0], 3);
            return -1;
        }
        adds++;
        longest = max(longest, w.len);
        if (w.len > 0) {
            bm_next_pivot(a, w);
            if (!(w.piv > p)) {
                if (fph_lane() == 0) atomicExch(&a.status[0], 3);
                return -1;
            }
        }
    }
    if (w.len == 0) {
        if (target >= 0) {  // zero here but paired in the complete reduction
            if (fph_lane() == 0) atomicExch(&a.status[0], 3);
            return -1;
        }
        return 0;
    }
    const long long at = bm_publish(a, w, al, c);
    if (at < 0) return -1;
    if (fph_lane() == 0) st_coh(&a.desc[c], (unsigned)at);
    const int len = w.len;
    bm_clear(w);
    return len;
}

// qs: [0] next column, [4] abort, [5] XCD taking part (-1: none yet).
// status: [0] error, [1] ran, [2] additions, [3] longest working column,
// [4..5] uint64 entries stored, [6] columns stored.
__global__ __launch_bounds__(kPhaseBWaves * FPH_WAVE) void k_replay(
    RedArgs a, const unsigned* __restrict__ order, const int* __restrict__ res_count,
    int* __restrict__ qs, long long spin_limit) {
    extern __shared__ unsigned smem[];
    __shared__ int s_part;
    // the workgroups on one XCD take part (the columns they publish stay in
    // its L2), as in the device-wide queue
    if (threadIdx.x == 0) {
        const int xcc = (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 0xF);
        const int lead = atomicCAS(qs + 5, -1, xcc);
        s_part = lead == -1 || lead == xcc;
    }
    __syncthreads();
    if (!s_part) return;
    const int wid = threadIdx.x / FPH_WAVE;
    const int lane = fph_lane();
    unsigned* arena = smem + wid * kArenaB;
    for (int i = lane; i < kArenaB; i += FPH_WAVE) arena[i] = 0u;
    wave_sync();
    const int R = *res_count;
    Alloc al = make_alloc(512);
    int adds = 0, longest = 0, nstored = 0;
    unsigned long long stored = 0;
    while (true) {
        if (__shfl(ld_coh(reinterpret_cast<const unsigned*>(qs + 4)), 0)) break;
        unsigned k = 0;
        if (lane == 0) k = atomicAdd(reinterpret_cast<unsigned*>(qs), 1u);
        k = __shfl(k, 0);
        if (k >= (unsigned)R) break;
        const int len = replay_column(a, (int)order[k], arena, al, qs, adds, longest, spin_limit);
        if (len < 0) {
            if (lane == 0) atomicExch(qs + 4, 1);
            break;
        }
        if (len > 0) {
            stored += (unsigned long long)len;
            nstored++;
        }
    }
    if (lane == 0) {
        if (adds) atomicAdd(&a.status[2], adds);
        if (longest) atomicMax(&a.status[3], longest);
        if (stored) atomicAdd(reinterpret_cast<unsigned long long*>(a.status + 4), stored);
        if (nstored) atomicAdd(&a.status[6], nstored);
    }
    if (threadIdx.x == 0) atomicMax(&a.status[1], 1);
}

// The apparent columns' share of the reference's store (it holds them from
// the start, _numba_utils.py:68-104): their entries and count, and their
// longest column.
__global__ void k_replay_apparent(const unsigned char* __restrict__ kind,
                                  const int* __restrict__ off, long long K,
                                  int* __restrict__ status) {
    __shared__ unsigned long long s_len;
    __shared__ int s_cnt, s_max;
    if (threadIdx.x == 0) {
        s_len = 0;
        s_cnt = 0;
        s_max = 0;
    }
    __syncthreads();
    unsigned long long len = 0;
    int cnt = 0, mx = 0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < K;
         c += (long long)gridDim.x * blockDim.x) {
        if (kind[c] == 2) {
            const int n = off[c + 1] - off[c];
            len += (unsigned long long)n;
            cnt++;
            mx = max(mx, n);
        }
    }
    if (len) atomicAdd(&s_len, len);
    if (cnt) atomicAdd(&s_cnt, cnt);
    if (mx) atomicMax(&s_max, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_len) atomicAdd(reinterpret_cast<unsigned long long*>(status + 4), s_len);
        if (s_cnt) atomicAdd(&status[6], s_cnt);
        if (s_max) atomicMax(&status[3], s_max);
    }
}

__global__ void k_replay_init(int* __restrict__ status, unsigned long long* __restrict__ pool_ptr,
                              int* __restrict__ qs) {
    if (threadIdx.x < 8) status[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        *pool_ptr = 0;
        qs[0] = 0;
        qs[4] = 0;
        qs[5] = -1;
    }
}

__global__ void k_fill_u32(unsigned* __restrict__ p, long long n, unsigned v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------
// The reference's own semantics, one wavefront per segment, in its order.
//
// _general_cohomology_reduce_v2 (cohomology_general.py:202-371) stores
// reduced columns in a pool of pool_cap entries (reduce_h1_gpu.py:286-289)
// and drops a column that does not fit: its pivot is still paired, but a
// later column reaching that pivot stops there instead of adding it.  With
// drops (and MAX_COL truncation, _numba_utils.py:10, 43-65) the pairs depend
// on the order of the work, so this kernel follows it exactly: apparent
// columns enter the store in ascending column order (_numba_utils.py:68-104),
// then residual columns are reduced one at a time in descending rank order.
// seg_pool[2 s] = word offset of segment s's store, seg_pool[2 s + 1] = its
// capacity in reference entries; a stored column takes len + 1 words here.
constexpr int kSerialMaxCol = 1048576;

// One workgroup of kSerThreads per segment.  Every column addition is done
// by the whole workgroup: the merged order of the two sorted lists is cut
// into kSerThreads pieces along the merge path (ties: working column first;
// a piece never separates equal rows), each thread counts the survivors of
// its piece, a workgroup prefix sum places them, and each thread writes its
// piece.  The working column lives in LDS (two halves, ping-pong) while it
// fits, the owner column is staged into LDS while it fits, and longer lists
// are merged in two reusable off-chip buffers.
constexpr int kSerThreads = 256;
constexpr int kSegPoolWords = 5;  // per segment: store base, capacity, row0, rows, bitmap
constexpr int kSerWaves = kSerThreads / FPH_WAVE;
constexpr int kSerHalf = 12288;   // words per LDS half of the working column
constexpr int kSerStage = 8192;   // words of LDS staging for the owner column
constexpr size_t kSerLds = (2 * kSerHalf + kSerStage) * 4;

struct SerShared {
    int split_i[kSerThreads + 1];
    int split_j[kSerThreads + 1];
    int wave_sum[kSerWaves];
    int flag;
    int first_bad;
    unsigned long long big_off[2];
    unsigned long long next_v;  // owner word of the result's pivot (0: none)
    int next_nb;                // that owner's length
    long long store_at;         // store offset when allocated from the pool
    int red[2][kSerWaves][3];   // bitmap form: per-wave (sum, bad, max), by parity
    int wfirst[2][kSerWaves];   // bitmap form: per-wave lowest set row, by parity
};

// Exclusive prefix sum of v over the workgroup; *total = sum.  Contains
// barriers: call from uniform control flow.
__device__ int block_excl_scan(int v, SerShared& sh, int* total) {
    const int lane = fph_lane(), wid = threadIdx.x / FPH_WAVE;
    int incl = v;
    for (int o = 1; o < FPH_WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == FPH_WAVE - 1) sh.wave_sum[wid] = incl;
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < kSerWaves; w++) {
        const int x = sh.wave_sum[w];
        if (w < wid) before += x;
        all += x;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

// C = A xor B (sorted, duplicate-free; any sizes), keeping at most `limit`
// rows (a prefix); C must not alias A or B.  Returns the length.  The
// owner of the result's pivot is looked up while the result is written:
// sh.next_v / sh.next_nb (valid when the result is not empty).
__device__ int block_xor(const unsigned* A, int na, const unsigned* B, int nb, unsigned* C,
                         int limit, SerShared& sh, const RedArgs& a) {
    const int t = threadIdx.x;
    const int total = na + nb;
    const int per = (total + kSerThreads - 1) / kSerThreads;
    const int k = min(t * per, total);
    int lo = max(0, k - nb), hi = min(k, na);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] <= B[k - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    int i0 = lo;
    const int j0 = k - lo;
    if (i0 > 0 && j0 < nb && A[i0 - 1] == B[j0]) i0--;
    sh.split_i[t] = i0;
    sh.split_j[t] = j0;
    if (t == 0) {
        sh.split_i[kSerThreads] = na;
        sh.split_j[kSerThreads] = nb;
    }
    __syncthreads();
    const int i1 = sh.split_i[t + 1], j1 = sh.split_j[t + 1];
    // Branch-free merge steps (the lanes of a wave take different branches
    // otherwise): both heads are compared, the smaller one survives unless
    // they are equal, and each list advances when its head was <= the other.
    constexpr unsigned kEnd = 0xFFFFFFFFu;
    int cnt = 0;
    unsigned first = kEnd;  // this piece's first survivor
    {
        int ia = i0, ib = j0;
        unsigned av = ia < i1 ? A[ia] : kEnd, bv = ib < j1 ? B[ib] : kEnd;
        while (av != kEnd || bv != kEnd) {
            const unsigned m = min(av, bv);
            const bool keep = av != bv;
            first = (keep && cnt == 0) ? m : first;
            cnt += keep;
            ia += av <= bv;
            ib += bv <= av;
            av = ia < i1 ? A[ia] : kEnd;
            bv = ib < j1 ? B[ib] : kEnd;
        }
    }
    // speculative: only the piece holding the result's first row keeps it
    const unsigned long long pv = first != kEnd ? a.own[(long long)first - a.row_base] : 0ull;
    int all;
    int at = block_excl_scan(cnt, sh, &all);
    const int len = min(all, limit);
    const bool lead = cnt > 0 && at == 0;
    int pnb = 0;
    if (lead && pv != kNoOwner) pnb = (int)a.pool[pv & 0xFFFFFFFFull];
    if (at < len) {
        int ia = i0, ib = j0;
        unsigned av = ia < i1 ? A[ia] : kEnd, bv = ib < j1 ? B[ib] : kEnd;
        while ((av != kEnd || bv != kEnd) && at < len) {
            const unsigned m = min(av, bv);
            if (av != bv) C[at++] = m;
            ia += av <= bv;
            ib += bv <= av;
            av = ia < i1 ? A[ia] : kEnd;
            bv = ib < j1 ? B[ib] : kEnd;
        }
    }
    if (lead) {
        sh.next_v = pv;
        sh.next_nb = pnb;
    }
    __threadfence_block();
    __syncthreads();
    return len;
}

__device__ __forceinline__ void block_copy(const unsigned* src, int n, unsigned* dst) {
    for (int i = threadIdx.x; i < n; i += kSerThreads) dst[i] = src[i];
}

// ---- working column as a bitmap --------------------------------------
//
// The segment's rows [row_lo, row_lo + rows) as a bitmap (in LDS when it
// fits next to its summary, else in a zeroed global region) plus a summary
// bitmap in LDS (bit w set iff bitmap word w is non-zero).  Adding a pooled
// column (sorted rows) flips one bit per row with independent atomics; the
// pivot (lowest set bit; it only increases) is found through the summary;
// at the end the column is read out in ascending order through the summary
// and the bits are cleared.  Same additions in the same order as the list
// form, so the same results.
constexpr int kBmLdsWords = 2 * kSerHalf + kSerStage;  // dynamic LDS words

// A global bitmap belongs to one workgroup (its segment's), so its atomics,
// loads and stores are workgroup-scoped: they stay in the L2 of the XCD the
// workgroup runs on instead of going device-wide.
template <bool G>
__device__ __forceinline__ unsigned bm_load(const unsigned* p) {
    if constexpr (G) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else return *p;
}

template <bool G>
__device__ __forceinline__ void bm_zero(unsigned* p) {
    if constexpr (G) __hip_atomic_store(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else *p = 0u;
}

// Flip local row x; returns +1 if the bit was set by it, -1 if cleared.
template <bool G>
__device__ __forceinline__ int bm_flip(unsigned* bits, unsigned* summ, int x) {
    const int w = x >> 5;
    const unsigned m = 1u << (x & 31);
    unsigned old;
    if constexpr (G)
        old = __hip_atomic_fetch_xor(&bits[w], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
        old = atomicXor(&bits[w], m);
    if (old == 0u || old == m) atomicXor(&summ[w >> 5], 1u << (w & 31));
    return (old & m) ? -1 : 1;
}

// Where a kept column of len rows goes: at base + stored in the segment's
// store, or -- base < 0, the complete reduction with no store capacity --
// bump-allocated from the pool.  -1 on pool overflow (status 1).  Contains
// a barrier: call from uniform control flow.
__device__ long long serial_store_at(const RedArgs& a, SerShared& sh, long long base,
                                     long long stored, int len) {
    if (base >= 0) return base + stored;
    if (threadIdx.x == 0) {
        const unsigned long long need = ((unsigned long long)len + 1 + kLine - 1) / kLine * kLine;
        const unsigned long long at = atomicAdd(a.pool_ptr, need);
        const bool over = at + need > a.pool_cap || at + need > 0xFFFFFFFFull;
        if (over) atomicExch(&a.status[0], 1);
        sh.store_at = over ? -1 : (long long)at;
    }
    __syncthreads();
    return sh.store_at;
}

template <bool G>
__device__ __forceinline__ void serial_bitmap(RedArgs& a, SerShared& sh, const unsigned* __restrict__ order,
                              int lo, int lo2, long long row_lo, int rows, unsigned* bits,
                              unsigned* summ, long long base, long long cap, long long& used,
                              long long& stored, int& nxor, int& longest, int maxcol) {
    const int t = threadIdx.x;
    const int S = ((rows + 31) / 32 + 31) / 32;  // summary words
    for (int r = lo; r < lo2; r++) {
        const int c = (int)order[r];
        const int lo_c = a.off[c];
        const int n0 = min(a.off[c + 1] - lo_c, maxcol);
        int bad = 0;
        for (int e = t; e < n0; e += kSerThreads) {
            const long long x = (long long)a.data[lo_c + e] - row_lo;
            if (x < 0 || x >= rows) bad = 1;
            else bm_flip<G>(bits, summ, (int)x);
        }
        int len = n0;
        int p = (int)((long long)a.data[lo_c] - row_lo);
        int phigh = (int)((long long)a.data[lo_c + n0 - 1] - row_lo);
        int nbad;
        (void)block_excl_scan(bad, sh, &nbad);  // also the barrier after the flips
        if (nbad) {
            if (t == 0) atomicExch(&a.status[0], 3);
            return;
        }
        // Each addition: flips, one barrier for the combined (sum, bad, max)
        // reduction, one for the lowest-set-bit search.  LDS slots
        // alternate by parity, so no barrier is needed to reuse them.
        int par = 0;
        const int wid = t / FPH_WAVE, lane = fph_lane();
        while (true) {
            const unsigned long long v = a.own[row_lo + p - a.row_base];
            if (v == kNoOwner) break;
            const unsigned long long vo = v & 0xFFFFFFFFull;
            const unsigned* ob = a.pool + vo;
            // the length and each thread's first row in one round trip (a
            // row past the column's end is a pool word, then ignored)
            const int nb = (int)ob[0];
            const unsigned r0 = vo + 1 + t < a.pool_cap ? ob[1 + t] : 0u;

            // the wave's sums by ballot (wave-uniform): change in length =
            // 2 x bits set - rows flipped; rows ascend, so the wave's last
            // valid row is its largest
            int delta = 0, lmax = 0, badw = 0;
            for (int base = 0; base < nb; base += kSerThreads) {
                const int e = base + t;
                const bool need = e < nb;
                long long x = 0;
                bool ok = false, set = false;
                if (need) {
                    x = (long long)(base == 0 ? r0 : ob[1 + e]) - row_lo;
                    ok = x >= 0 && x < rows;
                    if (ok) set = bm_flip<G>(bits, summ, (int)x) > 0;
                }
                const unsigned long long mok = __ballot(ok);
                badw |= __ballot(need && !ok) != 0ull;
                delta += 2 * __popcll(__ballot(set)) - __popcll(mok);
                if (mok)
                    lmax = max(lmax, __builtin_amdgcn_readlane((int)x, 63 - __clzll((long long)mok)));
            }
            if (lane == 0) {
                sh.red[par][wid][0] = delta;
                sh.red[par][wid][1] = badw;
                sh.red[par][wid][2] = lmax;
            }
            __syncthreads();
            int dsum = 0, nbad = 0;
            for (int w2 = 0; w2 < kSerWaves; w2++) {
                dsum += sh.red[par][w2][0];
                nbad |= sh.red[par][w2][1];
                phigh = max(phigh, sh.red[par][w2][2]);
            }
            if (nbad) {  // a row outside the segment
                if (t == 0) atomicExch(&a.status[0], 3);
                return;
            }
            if (t == 0) count_merged(a, 2LL * len + nb + dsum);
            len += dsum;
            nxor++;
            if (len > maxcol) {  // the reference would truncate: list form
                if (t == 0) atomicExch(&a.status[0], 4);
                return;
            }
            if (len == 0) break;
            // lowest set bit at or above the old pivot's word and at or
            // below the highest row flipped: each wave walks its interleaved
            // 64-word chunks of the summary and stops at its first chunk
            // with a nonzero word (ballot); that word's lowest set row is
            // the wave's candidate, in the wave's own LDS slot.  The row,
            // not the word, goes through LDS: after the barrier no wave
            // reads the bitmap, which the next addition's flips may already
            // change.  (The summary mirrors the nonzero words exactly: each
            // zero <-> nonzero transition of a word is seen by one flip.)
            const int ws = p >> 5;
            const int s_lo = ws >> 5, s_hi = min(phigh >> 10, S - 1);
            int cand = 0x7FFFFFFF;
            for (int i0 = s_lo + wid * FPH_WAVE; i0 <= s_hi; i0 += kSerThreads) {
                const int i = i0 + lane;
                unsigned val = i <= s_hi ? summ[i] : 0u;
                if (i == s_lo) val &= ~0u << (ws & 31);
                const unsigned long long hit = __ballot(val != 0u);
                if (hit) {
                    const int l = __ffsll((long long)hit) - 1;
                    if (lane == l) {
                        const int w = i * 32 + __builtin_ctz(val);
                        const unsigned bw = bm_load<G>(&bits[w]);
                        if (bw) cand = w * 32 + __builtin_ctz(bw);
                    }
                    cand = __builtin_amdgcn_readlane(cand, l);
                    break;
                }
            }
            if (lane == 0) sh.wfirst[par][wid] = cand;
            __syncthreads();
            int found = sh.wfirst[par][0];
            for (int w2 = 1; w2 < kSerWaves; w2++) found = min(found, sh.wfirst[par][w2]);
            const int np = found != 0x7FFFFFFF ? found : -1;
            if (!(np > p)) {  // pivot must increase
                if (t == 0) atomicExch(&a.status[0], 3);
                return;
            }
            p = np;
            par ^= 1;
        }
        if (len == 0) {
            if (t == 0) a.pivot[c] = -1;
            __syncthreads();
            continue;
        }
        if (t == 0) a.pivot[c] = (int)(row_lo + p);
        longest = max(longest, len);
        const bool keep = used + len <= cap;
        const long long at = keep ? serial_store_at(a, sh, base, stored, len) : 0;
        if (at < 0) return;
        // read out (when kept) and clear, summary words [p / 1024, phigh / 1024]
        int written = 0;
        for (int sw0 = p >> 10; sw0 <= (phigh >> 10); sw0 += kSerThreads) {
            const int sw = sw0 + t;
            const unsigned val = (sw <= (phigh >> 10) && sw < S) ? summ[sw] : 0u;
            int cnt = 0;
            for (unsigned m = val; m; m &= m - 1) cnt += __popc(bm_load<G>(&bits[sw * 32 + __builtin_ctz(m)]));
            int tot;
            int off = block_excl_scan(cnt, sh, &tot) + written;
            for (unsigned m = val; m; m &= m - 1) {
                const int w = sw * 32 + __builtin_ctz(m);
                unsigned bw = bm_load<G>(&bits[w]);
                if (keep)
                    for (; bw; bw &= bw - 1)
                        a.pool[at + 1 + off++] = (unsigned)(row_lo + w * 32 + __builtin_ctz(bw));
                bm_zero<G>(&bits[w]);
            }
            if (val) summ[sw] = 0u;
            written += tot;
        }
        if (keep) {
            if (t == 0) a.pool[at] = (unsigned)len;
            __threadfence_block();
            __syncthreads();
            if (t == 0)
                a.own[row_lo + p - a.row_base] =
                    ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
            used += len;
            stored += len + 1;
        }
        if (written != len && t == 0) atomicExch(&a.status[0], 3);
        __threadfence_block();
        __syncthreads();
        if (written != len) return;
    }
}

// ---- working column over an LDS window of a global bitmap -------------
//
// For segments whose bitmap does not fit LDS.  The rows of ww bitmap words
// starting at the word of the column's first pivot (the window) are kept in
// an LDS bitmap with its summary, as the LDS form keeps all rows; a row past
// the window is flipped in the segment's global bitmap by an atomic whose
// result is not waited for, and its word is marked in an LDS bitmap of
// touched global words.  The pivot only increases and every added column
// starts at the pivot, so no row below the window is ever flipped.  An
// addition thus waits on two global reads (the owner word, then the owner
// column) instead of four (also the flips' old values and the pivot's word).
// When the window holds no set row, the workgroup waits for its global
// flips, finds the first nonzero marked word and moves the window there;
// none left means the column is zero.  The column length is kept as an
// upper bound (rows at the start + rows added); when that passes MAX_COL the
// set rows are counted.  Same additions in the same order as the list form,
// so the same results.
//
// Marks exist only for words past the window (a window's range is copied
// into LDS and unmarked when it moves there); all are cleared, with their
// global words zeroed, when the column ends.
constexpr int kWinMinWords = 1024;  // smallest window worth this form

struct Win {
    unsigned* ts;    // marks: bit g <=> global word g may be nonzero
    unsigned* wsum;  // window summary: bit j <=> wb[j] != 0
    unsigned* wb;    // window bitmap: word j holds rows 32 (wlo + j) ..
    unsigned* gbits; // the segment's global bitmap (gw words)
    int gw;          // global words
    int tw;          // mark words
    int ww;          // window words (a multiple of 1024)
    int wlo;         // first global word of the window
};

__device__ __forceinline__ unsigned ld_agent(const unsigned* p) {
    return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Flips local row x (32 wlo <= x < rows).
__device__ __forceinline__ void win_flip(const Win& W, int x) {
    const int g = x >> 5;
    const unsigned m = 1u << (x & 31);
    const int j = g - W.wlo;
    if (j < W.ww) {
        const unsigned old = atomicXor(&W.wb[j], m);
        if (old == 0u || old == m) atomicXor(&W.wsum[j >> 5], 1u << (j & 31));
    } else {
        (void)__hip_atomic_fetch_xor(&W.gbits[g], m, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
        atomicOr(&W.ts[g >> 5], 1u << (g & 31));
    }
}

// Up to 8 marked words of mark word tw (bits m, taken from the lowest and
// removed): their indices to g[], values to v[], all loads in flight at
// once.  Returns how many.
__device__ __forceinline__ int win_load8(const Win& W, int tw, unsigned& m, int* g, unsigned* v) {
    int k = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (m) {
            g[q] = tw * 32 + __builtin_ctz(m);
            m &= m - 1;
            v[q] = ld_agent(&W.gbits[g[q]]);
            k = q + 1;
        }
    }
    return k;
}

// Workgroup OR of `bad` and maximum of `hi`.  Barriers; uniform results.
__device__ void win_reduce(SerShared& sh, int bad, int hi, int* nbad, int* hmax) {
    const int wid = threadIdx.x / FPH_WAVE;
    const int b = __ballot(bad != 0) != 0ull;
    for (int o = FPH_WAVE / 2; o; o >>= 1) hi = max(hi, __shfl_xor(hi, o));
    if (fph_lane() == 0) {
        sh.red[0][wid][1] = b;
        sh.red[0][wid][2] = hi;
    }
    __syncthreads();
    int nb = 0, h = 0;
    for (int w = 0; w < kSerWaves; w++) {
        nb |= sh.red[0][w][1];
        h = max(h, sh.red[0][w][2]);
    }
    __syncthreads();
    *nbad = nb;
    *hmax = h;
}

// Lowest set row (local) of the window at or above window word j0, -1 if
// none, -2 if the summary and the bitmap disagree; found by the calling
// wavefront alone (every wavefront finds the same while no flips run).
__device__ __forceinline__ int win_search(const Win& W, int j0) {
    const int lane = fph_lane();
    const int S = W.ww >> 5;
    for (int i0 = j0 >> 5; i0 < S; i0 += FPH_WAVE) {
        const int i = i0 + lane;
        unsigned val = i < S ? W.wsum[i] : 0u;
        if (i == (j0 >> 5)) val &= ~0u << (j0 & 31);
        const unsigned long long hit = __ballot(val != 0u);
        if (hit) {
            const int l = __ffsll((long long)hit) - 1;
            const int j = __builtin_amdgcn_readlane(i * 32 + (val ? __builtin_ctz(val) : 0), l);
            const unsigned bw = W.wb[j];
            return bw ? (W.wlo + j) * 32 + __builtin_ctz(bw) : -2;
        }
    }
    return -1;
}

// The first nonzero marked global word at or past g0 and at most gmax, or
// -1; marks of the zero words passed on the way are cleared.  After the
// workgroup's global flips have completed.  Barriers; uniform result.
__device__ int win_next_word(Win& W, SerShared& sh, int g0, int gmax) {
    const int t = threadIdx.x, wid = t / FPH_WAVE;
    const int te = min(W.tw - 1, gmax >> 5);
    for (int tb = g0 >> 5; tb <= te; tb += kSerThreads) {
        const int tw = tb + t;
        unsigned bits = tw <= te ? W.ts[tw] : 0u;
        if (tw == (g0 >> 5)) bits &= ~0u << (g0 & 31);
        int first = 0x7FFFFFFF;
        unsigned zero = 0u;
        for (unsigned m = bits; m; m &= m - 1) {
            const int g = tw * 32 + __builtin_ctz(m);
            if (ld_agent(&W.gbits[g])) {
                first = g;
                break;
            }
            zero |= m & (~m + 1u);
        }
        if (zero) W.ts[tw] &= ~zero;
        for (int o = FPH_WAVE / 2; o; o >>= 1) first = min(first, __shfl_xor(first, o));
        if (fph_lane() == 0) sh.wave_sum[wid] = first;
        __syncthreads();
        int all = sh.wave_sum[0];
        for (int w = 1; w < kSerWaves; w++) all = min(all, sh.wave_sum[w]);
        __syncthreads();
        if (all != 0x7FFFFFFF) return all;
    }
    return -1;
}

// Moves the (empty) window to start at global word gn: the marked words of
// [gn, gn + ww) are copied into it, zeroed and unmarked.  Barriers.
__device__ void win_move(Win& W, int gn) {
    const int t = threadIdx.x;
    W.wlo = gn;
    const int ge = min(gn + W.ww, W.gw);
    for (int g = gn + t; g < ge; g += kSerThreads) {
        if ((W.ts[g >> 5] >> (g & 31)) & 1u) {
            const unsigned v = ld_agent(&W.gbits[g]);
            if (v) {
                const int j = g - gn;
                W.wb[j] = v;
                atomicOr(&W.wsum[j >> 5], 1u << (j & 31));
                __hip_atomic_store(&W.gbits[g], 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    __syncthreads();  // every mark read before any is cleared
    for (int tw = (gn >> 5) + t; tw <= (ge - 1) >> 5; tw += kSerThreads) {
        const int lo = max(gn - tw * 32, 0), hi = min(ge - tw * 32, 32);  // bits [lo, hi)
        const unsigned m = (hi == 32 ? ~0u : (1u << hi) - 1u) & (~0u << lo);
        W.ts[tw] &= ~m;
    }
    vm_drain();
    __syncthreads();
}

// Set rows of window summary words [s0, s1] and marked global words of mark
// words [m0, m1].  After the global flips have completed.  Barriers.
__device__ int win_count(const Win& W, SerShared& sh, int s0, int s1, int m0, int m1) {
    const int t = threadIdx.x;
    int cnt = 0;
    for (int sw = s0 + t; sw <= s1; sw += kSerThreads)
        for (unsigned m = W.wsum[sw]; m; m &= m - 1) cnt += __popc(W.wb[sw * 32 + __builtin_ctz(m)]);
    for (int tw = m0 + t; tw <= m1; tw += kSerThreads) {
        unsigned m = W.ts[tw];
        while (m) {
            int g[8];
            unsigned v[8];
            const int k = win_load8(W, tw, m, g, v);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < k) cnt += __popc(v[q]);
        }
    }
    int tot;
    (void)block_excl_scan(cnt, sh, &tot);
    return tot;
}

// Writes the column's rows in ascending order to dst (unless null) and
// clears it: window summary words [s0, s1] and marked global words of mark
// words [m0, m1] (zeroed, unmarked).  Returns the rows found.  After the
// global flips have completed.  Barriers.
__device__ int win_drain(Win& W, SerShared& sh, int s0, int s1, int m0, int m1, unsigned* dst,
                         long long row_lo) {
    const int t = threadIdx.x;
    int written = 0;
    for (int sb = s0; sb <= s1; sb += kSerThreads) {
        const int sw = sb + t;
        const unsigned val = sw <= s1 ? W.wsum[sw] : 0u;
        int cnt = 0;
        for (unsigned m = val; m; m &= m - 1) cnt += __popc(W.wb[sw * 32 + __builtin_ctz(m)]);
        int tot;
        int at = written + block_excl_scan(cnt, sh, &tot);
        for (unsigned m = val; m; m &= m - 1) {
            const int j = sw * 32 + __builtin_ctz(m);
            unsigned bw = W.wb[j];
            if (dst)
                for (; bw; bw &= bw - 1)
                    dst[at++] = (unsigned)(row_lo + (long long)(W.wlo + j) * 32 + __builtin_ctz(bw));
            W.wb[j] = 0u;
        }
        if (val) W.wsum[sw] = 0u;
        written += tot;
    }
    for (int tb = m0; tb <= m1; tb += kSerThreads) {
        const int tw = tb + t;
        const unsigned bits = tw <= m1 ? W.ts[tw] : 0u;
        int cnt = 0;
        for (unsigned m = bits; m;) {
            int g[8];
            unsigned v[8];
            const int k = win_load8(W, tw, m, g, v);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < k) cnt += __popc(v[q]);
        }
        int tot;
        int at = written + block_excl_scan(cnt, sh, &tot);
        for (unsigned m = bits; m;) {
            int g[8];
            unsigned v[8];
            const int k = win_load8(W, tw, m, g, v);
            for (int q = 0; q < k; q++) {
                if (dst)
                    for (unsigned bw = v[q]; bw; bw &= bw - 1)
                        dst[at++] = (unsigned)(row_lo + (long long)g[q] * 32 + __builtin_ctz(bw));
                __hip_atomic_store(&W.gbits[g[q]], 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (bits) W.ts[tw] = 0u;
        written += tot;
    }
    vm_drain();
    __syncthreads();
    return written;
}

__device__ void serial_window(RedArgs& a, SerShared& sh, const unsigned* __restrict__ order,
                              int lo, int lo2, long long row_lo, int rows, unsigned* gbits,
                              unsigned* smem, int ww, long long base, long long cap,
                              long long& used, long long& stored, int& nxor, int& longest,
                              int maxcol) {
    const int t = threadIdx.x;
    Win W;
    W.gw = (rows + 31) / 32;
    W.tw = (W.gw + 31) / 32;
    W.ww = ww;
    W.ts = smem;
    W.wsum = smem + W.tw;
    W.wb = smem + W.tw + (W.ww >> 5);
    W.gbits = gbits;
    W.wlo = 0;
    long long merged = 0;
    for (int r = lo; r < lo2; r++) {
        const int c = (int)order[r];
        const int lo_c = a.off[c];
        const int n0 = min(a.off[c + 1] - lo_c, maxcol);
        const long long p0 = (long long)a.data[lo_c] - row_lo;
        if (p0 < 0 || p0 >= rows) {  // uniform
            if (t == 0) atomicExch(&a.status[0], 3);
            return;
        }
        W.wlo = (int)(p0 >> 5);
        int bad = 0, hi = (int)p0;  // per thread: a row outside, highest row flipped
        for (int e = t; e < n0; e += kSerThreads) {
            const long long x = (long long)a.data[lo_c + e] - row_lo;
            if (x < 32LL * W.wlo || x >= rows) bad = 1;
            else {
                win_flip(W, (int)x);
                hi = max(hi, (int)x);
            }
        }
        merged += n0;
        long long ub = n0;  // >= the column's length
        int p = (int)p0;
        bool zero = false;
        // Per addition: the owner read, a barrier (every wavefront's search
        // of the previous pivot is done; it waits while the read is in
        // flight), the flips, a barrier (the flips are visible), and the
        // search for the new pivot by every wavefront on its own.
        while (true) {
            const unsigned long long v = a.own[row_lo + p - a.row_base];
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (v == kNoOwner) break;
            const unsigned long long vo = v & 0xFFFFFFFFull;
            const unsigned* ob = a.pool + vo;
            // the length and each thread's first row in one round trip (a
            // row past the column's end is a pool word, then ignored)
            const int nb = (int)ob[0];
            const unsigned r0 = vo + 1 + t < a.pool_cap ? ob[1 + t] : 0u;
            for (int b = 0; b < nb; b += kSerThreads) {
                const int e = b + t;
                if (e < nb) {
                    const long long x = (long long)(b == 0 ? r0 : ob[1 + e]) - row_lo;
                    if (x < 32LL * W.wlo || x >= rows) bad = 1;
                    else {
                        win_flip(W, (int)x);
                        hi = max(hi, (int)x);
                    }
                }
            }
            merged += nb;
            nxor++;
            ub += nb;
            if (ub > maxcol) {  // uniform: count the rows
                vm_drain();
                __syncthreads();
                int nbad, hmax;
                win_reduce(sh, bad, hi, &nbad, &hmax);
                const int n = win_count(W, sh, 0, (W.ww >> 5) - 1, (W.wlo + W.ww) >> 5,
                                        min(W.tw - 1, hmax >> 10));
                if (n > maxcol) {  // the reference would truncate: list form
                    if (t == 0) atomicExch(&a.status[0], 4);
                    return;
                }
                ub = n;
            }
            __syncthreads();  // the flips are visible
            int np = win_search(W, (p >> 5) - W.wlo);
            if (np == -1) {  // the window is empty: move it to the next set row
                vm_drain();
                __syncthreads();
                int nbad, hmax;
                win_reduce(sh, bad, hi, &nbad, &hmax);
                const int gn = win_next_word(W, sh, W.wlo + W.ww, hmax >> 5);
                if (gn < 0) {
                    zero = true;
                    break;
                }
                win_move(W, gn);
                const unsigned w0 = W.wb[0];  // gn's word, nonzero
                np = w0 ? gn * 32 + __builtin_ctz(w0) : -1;
            }
            if (!(np > p)) {  // pivot must increase
                if (t == 0) atomicExch(&a.status[0], 3);
                return;
            }
            p = np;
        }
        vm_drain();
        int nbad, hmax;
        win_reduce(sh, bad, hi, &nbad, &hmax);  // also the barrier after the flips
        if (nbad) {  // a row outside the segment
            if (t == 0) atomicExch(&a.status[0], 3);
            return;
        }
        if (zero) {  // window empty, every marked word zero and unmarked
            if (t == 0) a.pivot[c] = -1;
            continue;
        }
        if (t == 0) a.pivot[c] = (int)(row_lo + p);
        const int s0 = ((p >> 5) - W.wlo) >> 5;
        const int s1 = min((W.ww >> 5) - 1, ((hmax >> 5) - W.wlo) >> 5);
        const int m0 = (W.wlo + W.ww) >> 5, m1 = min(W.tw - 1, hmax >> 10);
        int len, written;
        bool keep;
        long long at;
        if (base >= 0 && used + ub <= cap) {
            // the segment's store takes the column whatever its length: read
            // it out once, straight into place
            keep = true;
            at = base + stored;
            len = written = win_drain(W, sh, s0, s1, m0, m1, a.pool + at + 1, row_lo);
        } else {
            len = win_count(W, sh, s0, s1, m0, m1);
            keep = used + len <= cap;
            at = keep ? serial_store_at(a, sh, base, stored, len) : 0;
            if (at < 0) return;
            written = win_drain(W, sh, s0, s1, m0, m1, keep ? a.pool + at + 1 : nullptr, row_lo);
        }
        longest = max(longest, len);
        if (keep) {
            if (t == 0) a.pool[at] = (unsigned)len;
            __threadfence_block();
            __syncthreads();
            if (t == 0)
                a.own[row_lo + p - a.row_base] =
                    ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
            used += len;
            stored += len + 1;
        }
        if (written != len && t == 0) atomicExch(&a.status[0], 3);
        __threadfence_block();
        __syncthreads();
        if (written != len) return;
    }
    if (t == 0 && merged) count_merged(a, merged);
}

// The reference's pool prefill (_numba_utils.py:68-104, reduce_h1_gpu.py:
// 282-289): apparent columns of [c0, c1) in ascending column order enter the
// store while it has room (cap reference entries); a column that does not
// fit is dropped and later ones may still fit.  Sets own[] for the stored
// ones (their rows in place in the pool, written by k_publish_apparent) and
// returns the entries used.  Whole workgroup of kSerThreads; barriers.
__device__ long long serial_prefill(const RedArgs& a, SerShared& sh, int c0, int c1,
                                    long long cap) {
    const int t = threadIdx.x;
    long long used = 0;
    bool fast = true;  // until the store first fills up
    for (int cb = c0; cb < c1;) {
        if (fast) {
            // four chunks at once (kinds and offsets in one round trip, first
            // rows in another) when all of them fit
            bool app4[4];
            int cn4[4], lo4[4], sum = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int c = cb + k * kSerThreads + t;
                const bool in = c < c1;
                const unsigned char kd = in ? a.kind[c] : 0;
                const int lo = in ? a.off[c] : 0, hi = in ? a.off[c + 1] : 0;
                app4[k] = in && kd == 2;
                cn4[k] = app4[k] ? hi - lo : 0;
                lo4[k] = lo;
                sum += cn4[k];
            }
            int total;
            (void)block_excl_scan(sum, sh, &total);
            if (used + total <= cap) {
                unsigned first[4];
#pragma unroll
                for (int k = 0; k < 4; k++) first[k] = cn4[k] > 0 ? a.data[lo4[k]] : 0u;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (cn4[k] > 0) {
                        const int c = cb + k * kSerThreads + t;
                        a.own[(long long)first[k] - a.row_base] =
                            ((unsigned long long)(c + 1) << 32) | (unsigned long long)(lo4[k] + c);
                    }
                used += total;
                cb += 4 * kSerThreads;
                continue;
            }
            fast = false;
        }
        const int c = cb + t;
        cb += kSerThreads;
        const bool app = c < c1 && a.kind[c] == 2;
        const int cn = app ? a.off[c + 1] - a.off[c] : 0;
        int chunk;
        const int excl = block_excl_scan(cn, sh, &chunk);
        const int incl = excl + cn;
        if (used + chunk <= cap) {  // common case: the whole chunk fits
            if (app && cn > 0) {
                const long long lo = a.off[c];
                a.own[(long long)a.data[lo] - a.row_base] =
                    ((unsigned long long)(c + 1) << 32) | (unsigned long long)(lo + c);
            }
            used += chunk;
            continue;
        }
        // some column fails: walk the chunk, storing maximal fitting runs
        int p = 0;
        while (p < kSerThreads) {
            if (t == 0) sh.first_bad = kSerThreads;
            __syncthreads();
            // excl of thread p: entries of threads [0, p) in this chunk
            if (t == p) sh.flag = excl;
            __syncthreads();
            const int excl_p = sh.flag;
            if (t >= p && cn > 0 && used + (incl - excl_p) > cap) atomicMin(&sh.first_bad, t);
            __syncthreads();
            const int f = sh.first_bad;
            if (t >= p && t < f && cn > 0) {
                const long long lo = a.off[c];
                a.own[(long long)a.data[lo] - a.row_base] =
                    ((unsigned long long)(c + 1) << 32) | (unsigned long long)(lo + c);
            }
            if (t == f - 1) sh.flag = incl;
            __syncthreads();
            if (f > p) used += sh.flag - excl_p;
            __syncthreads();
            p = f + 1;  // column f (if any) is dropped
        }
    }
    return used;
}

__global__ __launch_bounds__(kSerThreads) void k_reduce_serial(
    RedArgs a, const unsigned* __restrict__ order, const int* __restrict__ res_count,
    const int* __restrict__ seg_cols, const long long* __restrict__ seg_pool, int variant,
    int maxcol_ref) {
    extern __shared__ unsigned smem[];
    __shared__ SerShared sh;
    unsigned* half0 = smem;
    unsigned* stage = smem + 2 * kSerHalf;
    const int t = threadIdx.x;
    const int seg = blockIdx.x;
    const int c0 = seg_cols ? seg_cols[seg] : 0;
    const int c1 = seg_cols ? seg_cols[seg + 1] : (int)a.K;
    const long long* sp = seg_pool + kSegPoolWords * seg;
    const long long base = sp[0];
    const long long cap = sp[1];
    long long stored = 0;  // words written at base (uniform)
    // 1. apparent columns, ascending: store while pool_ptr + cn <= pool_cap
    long long used = serial_prefill(a, sh, c0, c1, cap);  // reference entries in the store
    __threadfence_block();
    __syncthreads();
    // 2. residual columns of the segment, descending rank
    const int R = *res_count;
    int lo = 0, hi = R;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)order[mid] >= c1) lo = mid + 1;
        else hi = mid;
    }
    int lo2 = lo, hi2 = R;
    while (lo2 < hi2) {
        const int mid = (lo2 + hi2) >> 1;
        if ((int)order[mid] >= c0) lo2 = mid + 1;
        else hi2 = mid;
    }
    int nxor = 0, longest = 0, offchip = 0;
    // bitmap form: sp[2] = first row (global rank), sp[3] = rows, sp[4] =
    // word offset of a zeroed global bitmap in the pool, or -1 for LDS
    const int rows = (int)sp[3];
    // variant bit 0: sorted-list working column; bit 1: the complete
    // reduction (no store capacity, no MAX_COL truncation); bit 3: a global
    // bitmap without the LDS window; bits 8..15: at most that many 1024-word
    // units of window (0: as many as LDS holds)
    const int maxcol = (variant & 2) ? 0x7FFFFFFF : maxcol_ref;
    if ((variant & 1) == 0 && rows > 0) {
        const int W = (rows + 31) / 32, S = (W + 31) / 32;
        int ww = ((kBmLdsWords - S) * 32 / 33) & ~1023;  // window words beside the marks
        if ((variant >> 8) & 0xFF) ww = min(ww, ((variant >> 8) & 0xFF) * 1024);
        const bool window = sp[4] >= 0 && (variant & 8) == 0 && ww >= kWinMinWords;
        const int zw = window ? kBmLdsWords : S + (sp[4] < 0 ? W : 0);
        for (int i = t; i < zw; i += kSerThreads) smem[i] = 0u;
        __syncthreads();
        if (sp[4] < 0)
            serial_bitmap<false>(a, sh, order, lo, lo2, sp[2], rows, smem + S, smem, base, cap,
                                 used, stored, nxor, longest, maxcol);
        else if (window)
            serial_window(a, sh, order, lo, lo2, sp[2], rows, a.pool + sp[4], smem, ww, base,
                          cap, used, stored, nxor, longest, maxcol);
        else
            serial_bitmap<true>(a, sh, order, lo, lo2, sp[2], rows, a.pool + sp[4], smem, base,
                                cap, used, stored, nxor, longest, maxcol);
        lo = lo2;  // done (or aborted with a status)
    }
    unsigned long long big_off[2] = {0, 0};
    long long big_cap[2] = {0, 0};
    for (int r = lo; r < lo2; r++) {
        const int c = (int)order[r];
        const int lo_c = a.off[c];
        const int n0 = min(a.off[c + 1] - lo_c, maxcol);
        const unsigned* cur;
        int where;  // 0 / 1: LDS half, 2 / 3: off-chip buffer, -1: CSR (read-only)
        if (n0 <= kSerHalf) {
            block_copy(a.data + lo_c, n0, half0);
            cur = half0;
            where = 0;
        } else {
            cur = a.data + lo_c;
            where = -1;
        }
        int len = n0;
        __syncthreads();
        bool have_next = false;
        unsigned long long next_v = 0;
        int next_nb = 0;
        while (len > 0) {
            const unsigned p_old = cur[0];
            unsigned long long v;
            int nb;
            if (have_next) {
                v = next_v;
                nb = next_nb;
            } else {
                v = a.own[(long long)p_old - a.row_base];
                nb = v != kNoOwner ? (int)a.pool[v & 0xFFFFFFFFull] : 0;
            }
            if (v == kNoOwner) break;
            const unsigned* ob = a.pool + (v & 0xFFFFFFFFull);
            const unsigned* B = ob + 1;
            if (nb <= kSerStage) {
                block_copy(ob + 1, nb, stage);
                B = stage;
            }
            const int need = len + nb;
            unsigned* out;
            int out_where;
            if ((where == 0 || where == 1) && need <= kSerHalf) {
                out_where = 1 - where;
                out = smem + out_where * kSerHalf;
            } else {
                offchip++;
                const int bi = where == 2 ? 1 : 0;
                if (big_cap[bi] < need) {
                    long long ncap = 16384;
                    while (ncap < need) ncap *= 2;
                    if (t == 0) {
                        const unsigned long long at = atomicAdd(a.pool_ptr, (unsigned long long)ncap);
                        sh.big_off[bi] = at;
                        sh.flag = (at + ncap > a.pool_cap) ? 1 : 0;
                        if (sh.flag) atomicExch(&a.status[0], 1);
                    }
                    __syncthreads();
                    if (sh.flag) return;  // uniform: every thread read the same flag
                    big_off[bi] = sh.big_off[bi];
                    big_cap[bi] = ncap;
                }
                out = a.pool + big_off[bi];
                out_where = 2 + bi;
            }
            __syncthreads();  // staged owner visible
            int nl = block_xor(cur, len, B, nb, out, maxcol, sh, a);
            next_v = sh.next_v;
            next_nb = sh.next_nb;
            have_next = true;
            if (out_where >= 2 && nl <= kSerHalf) {  // back on chip
                const int h = (where == 0 || where == 1) ? 1 - where : 0;
                block_copy(out, nl, smem + h * kSerHalf);
                __syncthreads();
                out = smem + h * kSerHalf;
                out_where = h;
            }
            if (t == 0) count_merged(a, (long long)len + nb + nl);
            cur = out;
            where = out_where;
            len = nl;
            nxor++;
            if (len > 0 && !(cur[0] > p_old)) {
                if (t == 0) atomicExch(&a.status[0], 3);
                return;
            }
        }
        if (len == 0) {
            if (t == 0) a.pivot[c] = -1;
            continue;
        }
        const unsigned piv = cur[0];
        if (t == 0) a.pivot[c] = (int)piv;
        longest = max(longest, len);
        if (used + len <= cap) {
            const long long at = serial_store_at(a, sh, base, stored, len);
            if (at < 0) return;
            if (t == 0) a.pool[at] = (unsigned)len;
            block_copy(cur, len, a.pool + at + 1);
            __threadfence_block();
            __syncthreads();
            if (t == 0)
                a.own[(long long)piv - a.row_base] =
                    ((unsigned long long)(c + 1) << 32) | (unsigned long long)at;
            used += len;
            stored += len + 1;
        }
        __threadfence_block();
        __syncthreads();
    }
    if (t == 0) {
        if (nxor) atomicAdd(&a.status[2], nxor);
        if (offchip) atomicAdd(&a.status[4], offchip);
        if (longest) atomicMax(&a.status[3], longest);
        atomicMax(&a.status[1], 1);
    }
}

// status[5] = pool words handed out (in 32-word lines, saturating): what a
// retry after a pool overflow needs at least.
__global__ void k_reduce_fini(int* __restrict__ status,
                              const unsigned long long* __restrict__ pool_ptr) {
    if (threadIdx.x == 0) {
        const unsigned long long lines = pool_ptr[0] / kLine;
        status[5] = lines > 0x7FFFFFFFull ? 0x7FFFFFFF : (int)lines;
    }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
    if (threadIdx.x == 0) p[0] = v;
}

__global__ void k_reduce_init(int* __restrict__ status, unsigned long long* __restrict__ pool_ptr,
                              const int* __restrict__ off, long long K, RedArgs a) {
    if (threadIdx.x == 0) {
        status[0] = 0;
        status[1] = 0;
        status[2] = 0;
        status[3] = 0;
        status[4] = 0;
        status[5] = 0;
        status[6] = 0;
        status[7] = 0;
        // apparent region: off[K] + K words, one line of padding (speculative
        // reads of the last apparent column), then line-aligned dynamic space
        pool_ptr[0] = ((unsigned long long)off[K] + K + kLine - 1) / kLine * kLine + kLine;
        // owner lines, when used, come next
        if (lines_of(a)) pool_ptr[0] += 2ull * kLineWords * (unsigned long long)a.nrows;
    }
}

// Sort keys: the residual columns, the unused tail set to 0 (0 sorts last
// in descending order; only the first res_count sorted keys are read, and a
// residual column 0 is equal to the padding).
__global__ void k_order_keys(const int* __restrict__ res_list, const int* __restrict__ res_count,
                             long long K, unsigned* __restrict__ keys) {
    long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < K) keys[r] = r < *res_count ? (unsigned)res_list[r] : 0u;
}

// order[0..R) = the residual columns (kind 3) in descending index, by one
// workgroup: thread t counts its run of columns (t = 0 the highest run), a
// workgroup scan gives each run its place, and the run is written top down.
// Replaces a descending radix sort of the residual list (several launches
// of block sort and merge passes at these sizes).
constexpr int kOrdThreads = 1024;
constexpr long long kOrderCompactMax = 1LL << 21;  // beyond: the radix sort

__global__ __launch_bounds__(kOrdThreads) void k_order_desc(const unsigned char* __restrict__ kind,
                                                            long long K,
                                                            unsigned* __restrict__ order) {
    __shared__ int s_wave[kOrdThreads / FPH_WAVE];
    const int t = threadIdx.x, lane = fph_lane(), wid = t / FPH_WAVE;
    const long long chunk = (K + kOrdThreads - 1) / kOrdThreads;
    const long long hi = K - (long long)t * chunk;
    const long long lo = hi - chunk > 0 ? hi - chunk : 0;
    int cnt = 0;
    for (long long c = lo; c < hi; c++) cnt += kind[c] == 3;
    int incl = cnt;
    for (int o = 1; o < FPH_WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == FPH_WAVE - 1) s_wave[wid] = incl;
    __syncthreads();
    int pos = incl - cnt;
    for (int w = 0; w < wid; w++) pos += s_wave[w];
    for (long long c = hi - 1; c >= lo; c--)
        if (kind[c] == 3) order[pos++] = (unsigned)c;
}

// ---------------------------------------------------------------------
// The reference's semantics with many columns in flight.
//
// In the reference's loop (cohomology_general.py:255-366) a column adds, at
// each pivot, the column stored for that pivot and stops at the first pivot
// with none; it is then stored iff it fits (cohomology_general.py:340-356).
// A stored column never changes and a pivot is stored at most once (a column
// that finds one stored keeps reducing), so every addition a column makes is
// the one the serial loop makes, whenever it is made.  Only the decision to
// stop depends on the order: a column may stop at pivot p only once every
// earlier column (in the reference's order) has settled, and whether it is
// stored depends on the store's fill after them.  So each wavefront takes
// the next column of a segment, reduces it as far as the stored columns
// allow, stages the result in the pool and publishes it as a candidate, and
// goes on to another column without waiting: candidates are settled in
// order by settling runs (pipe_hold), which store one by pointing own[p] at
// its staged rows, drop one that does not fit, or withdraw one whose pivot
// was stored meanwhile and hand it back to the run's wavefront, which adds
// on from the staged rows.  A run starts at a position whose predecessor
// has a record (the store's fill after it), taken by that position's
// publisher or by the run that reached it, whichever claims it.
//
// Coherence: all wavefronts of a segment run on one XCD (an XCD claims a
// whole segment, and a wavefront only takes columns of its XCD's segment),
// so every shared word (own[], records, stored rows) meets in that XCD's L2:
// loads are device-scope (they bypass the CU's L1), and a store that another
// one publishes is drained (s_waitcnt vmcnt(0)) before the publishing store
// is issued.

constexpr int kPipeLdsMax = 20224;                   // LDS words per wavefront at most (2 per CU)
constexpr unsigned long long kNoRecord = ~0ull;
constexpr int kPipeSlots = 16;                       // open segments per XCD
constexpr int kPipeNext = 8 * kPipeSlots;            // ctl: next segment to open
constexpr int kPipeAbort = kPipeNext + 1;            // ctl: stop (an error)
constexpr int kPipeCtl = kPipeNext + 8;              // ctl: [8 x kPipeSlots slots, next, abort]

struct PipeArgs {
    RedArgs a;
    const unsigned* order;          // residual columns, descending
    const long long* seg_pool;      // kSegPoolWords per segment
    const int* seg_lo;              // nseg: first position of each segment in order
    const int* seg_hi;              // nseg: one past its last position
    int* seg_next;                  // nseg: next position handed out, relative to seg_lo
    int* seg_done;                  // nseg: positions settled
    const int* seg_order;           // nseg: segments, most residual columns first
    const unsigned long long* seg_state;  // nseg: store fill after the prefill
    unsigned long long* rec;        // per position: store fill after it (used << 32 | words),
                                    // written where a settling run ends
    unsigned long long* fin;        // per position: candidate final pivot and length
    unsigned* stg;                  // per position: pool offset of its staged candidate
    int* claim;                     // per position: 2 once its own wavefront settles from it
    int* ctl;                       // kPipeCtl control words
    unsigned* gbits;                // global bitmaps, wmax words per workgroup (or null)
    long long wmax;
    unsigned long long chunk;       // pool words a wavefront reserves at a time for staging
    long long spin_limit;           // wall-clock ticks a wavefront may find no work
    int window;                     // global bitmaps get an LDS window (pipe_column_win)
    int win_cap;                    // at most this many 1024-word window units (0: no cap)
    int nseg;
    int arena;                      // LDS words per wavefront
    int maxcol;
    int exact;                      // the complete reduction: every column is stored
    int force_global;               // bitmaps in global memory even where LDS holds them
    int inflight;                   // columns of a segment handed out and not settled, at most
    int inflight_lds;               // the same for segments whose bitmaps fit the LDS arena
};

__device__ __forceinline__ unsigned long long ld_coh64(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_coh64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int ld_coh_i(const int* p) {
    return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// First error wins; every wavefront leaves at its next check of the abort word.
__device__ bool pipe_fail(const PipeArgs& P, int code) {
    if (fph_lane() == 0) {
        atomicCAS(&P.a.status[0], 0, code);
        atomicExch(&P.ctl[kPipeAbort], 1);
    }
    return false;
}

// Exclusive prefix sum over the wavefront (DPP: within rows of 16 lanes,
// then across rows); *total = the sum.
__device__ __forceinline__ int wave_excl_scan(int v, int* total) {
    int incl = v;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, false);  // row_shr:1
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, false);  // row_shr:2
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, false);  // row_shr:4
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, false);  // row_shr:8
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x142, 0xa, 0xf, false);  // row_bcast:15
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x143, 0xc, 0xf, false);  // row_bcast:31
    *total = __builtin_amdgcn_readlane(incl, FPH_WAVE - 1);
    return incl - v;
}

// The store's fill after the position before r (after the prefill for the
// segment's first position); kNoRecord while that position is unsettled.
__device__ __forceinline__ unsigned long long pipe_prev(const PipeArgs& P, int s, long long r) {
    return r == P.seg_lo[s] ? P.seg_state[s] : ld_coh64(&P.rec[r - 1]);
}

// Flips rows src[b0 + u * 64 + lane] (u < kPipeUnroll, < n) of a sorted
// column: all loads, then all bit flips, then the summary updates, so that
// the memory round trips of the chunks overlap.  delta += change in the
// working column's length, lmax = max(lmax, highest row flipped).  False
// when a row lies outside the segment.
constexpr int kPipeUnroll = 4;

template <bool G>
__device__ __forceinline__ bool pipe_flip_chunk(const unsigned* src, int n, int b0,
                                                long long row_lo, int rows, unsigned* bits,
                                                unsigned* summ, int& delta, int& lmax) {
    const int lane = fph_lane();
    int x[kPipeUnroll];
    bool ok[kPipeUnroll];
    unsigned old[kPipeUnroll];
    bool bad = false;
#pragma unroll
    for (int u = 0; u < kPipeUnroll; u++) {
        const int e = b0 + u * FPH_WAVE + lane;
        ok[u] = false;
        x[u] = 0;
        if (e < n) {
            const long long xx = (long long)ld_coh(src + e) - row_lo;
            ok[u] = xx >= 0 && xx < rows;
            bad |= !ok[u];
            x[u] = ok[u] ? (int)xx : 0;
        }
    }
#pragma unroll
    for (int u = 0; u < kPipeUnroll; u++) {
        old[u] = 0u;
        if (ok[u]) {
            const unsigned m = 1u << (x[u] & 31);
            if constexpr (G)
                old[u] = __hip_atomic_fetch_xor(&bits[x[u] >> 5], m, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
            else
                old[u] = atomicXor(&bits[x[u] >> 5], m);
        }
    }
#pragma unroll
    for (int u = 0; u < kPipeUnroll; u++) {
        const unsigned m = 1u << (x[u] & 31);
        if (ok[u] && (old[u] == 0u || old[u] == m)) {
            const int w = x[u] >> 5;
            atomicXor(&summ[w >> 5], 1u << (w & 31));
        }
        const unsigned long long mok = __ballot(ok[u]);
        delta += 2 * __popcll(__ballot(ok[u] && !(old[u] & m))) - __popcll(mok);
        if (mok) lmax = max(lmax, __builtin_amdgcn_readlane(x[u], 63 - __clzll((long long)mok)));
    }
    return !__ballot(bad);
}

// A column's candidate final pivot, published for the settling run:
// (p << 32) | len, len = 0 for a column reduced to zero; kNoRecord before.
// Its rows are staged in the pool first ([len, rows...] at stg[position]),
// so that settling can store it by pointing own[p] there.

// Settles positions from q (claimed by the caller; q's candidate is
// published) onwards while they are ready: one round trip reads the next 64
// candidates, their staging offsets and the owners of their pivots; then,
// in order, a candidate whose pivot has an owner (stored before the run or
// by an earlier candidate of it) must continue, which ends the run; any
// other is stored if it fits (own[p] set to its staged rows) and gets its
// pivot.  A run that reaches an unpublished position writes the record of
// its last settled one, then looks at that position again: published in the
// meantime, it claims it and goes on (its wavefront, having published, may
// have seen no record yet; one of the two sees the other's write).  Returns
// -1 when the run ended, or the position of the candidate that must
// continue, withdrawn and claimed, for the caller to take over (it settles
// from there itself once it publishes again).  state = used << 32 | words
// stored.
__device__ long long pipe_hold(const PipeArgs& P, int s, long long q, unsigned long long state) {
    const RedArgs& a = P.a;
    const int lane = fph_lane();
    const long long* sp = P.seg_pool + kSegPoolWords * s;
    const long long cap = sp[1], row_lo = sp[2];
    unsigned long long* own = a.own + (row_lo - a.row_base);
    const long long hi = P.seg_hi[s];
    long long used = (long long)(state >> 32), stored = (long long)(state & 0xFFFFFFFFull);
    long long last = -1;  // last position this run settled
    while (true) {
        const long long pos = q + lane;
        const int m = (int)min<long long>(FPH_WAVE, hi - q);
        unsigned long long f = lane < m ? ld_coh64(&P.fin[pos]) : kNoRecord;
        const unsigned long long rdy = __ballot(f != kNoRecord);
        int m2 = (int)(__ffsll((long long)~rdy) - 1);  // ready prefix
        if (~rdy == 0ull) m2 = FPH_WAVE;
        m2 = min(m2, m);
        const int pl = (int)(f >> 32), ll = (int)(f & 0xFFFFFFFFull);
        unsigned long long o = 0;
        unsigned sg = 0;
        if (lane < m2 && ll > 0) {
            o = ld_coh64(&own[pl]);
            sg = ld_coh(&P.stg[pos]);
        }
        int stop = m2;  // first position not settled
        bool kept = false;
        for (int l = 0; l < m2; l++) {
            const int p_l = __shfl(pl, l), len_l = __shfl(ll, l);
            const unsigned long long o_l = __shfl(o, l);
            if (len_l == 0) continue;  // reduced to zero
            const bool dup = __ballot(lane < l && kept && pl == p_l) != 0ull;
            if (o_l != kNoOwner || dup) {
                stop = l;
                break;
            }
            bool keep = true;
            if (!P.exact) {
                keep = used + len_l <= cap;
                if (keep) {
                    used += len_l;
                    stored += len_l + 1;
                }
            }
            if (lane == l) kept = keep;
        }
        if (lane < stop) {
            const int c = (int)P.order[pos];
            if (kept) st_coh64(&own[pl], ((unsigned long long)(c + 1) << 32) | sg);
            a.pivot[c] = ll > 0 ? (int)(row_lo + pl) : -1;
        }
        vm_drain();  // the next round's owner reads see these owners
        const bool cont = stop < m2;  // the candidate at `stop` continues
        if (cont && lane == stop) {
            // withdrawn, and claimed for this wavefront, which takes it over
            // (its first publisher may still try to claim it: that fails)
            st_coh64(&P.fin[pos], kNoRecord);
            __hip_atomic_store(&P.claim[pos], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0 && stop > 0) atomicAdd(&P.seg_done[s], stop);
        if (stop > 0) last = q + stop - 1;
        if (stop < m || q + m >= hi) {
            // the run ends: the record of its last settled position
            vm_drain();
            if (lane == 0 && last >= 0)
                st_coh64(&P.rec[last],
                         ((unsigned long long)used << 32) | (unsigned long long)stored);
            if (cont) return q + stop;
            if (stop == m) return -1;  // the segment's end
            vm_drain();
            const long long u = q + stop;
            if (ld_coh64(&P.fin[u]) == kNoRecord) return -1;
            int won = 0;
            if (lane == 0) won = atomicCAS(&P.claim[u], 0, 2) == 0;
            if (!__shfl(won, 0)) return -1;
            q = u;
            continue;
        }
        q += m;
    }
}

// The first 63 rows of a column of nb rows whose words [len, rows...] lane l
// loaded into w (lane l > 0 holds row l - 1); as pipe_flip_chunk.
template <bool G>
__device__ __forceinline__ bool pipe_flip_first(unsigned w, int nb, long long row_lo, int rows,
                                                unsigned* bits, unsigned* summ, int& delta,
                                                int& lmax) {
    const int lane = fph_lane();
    const bool need = lane >= 1 && lane <= nb;
    const long long xx = (long long)w - row_lo;
    const bool ok = need && xx >= 0 && xx < rows;
    const int x = ok ? (int)xx : 0;
    const unsigned m = 1u << (x & 31);
    unsigned old = 0u;
    if (ok) {
        if constexpr (G)
            old = __hip_atomic_fetch_xor(&bits[x >> 5], m, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            old = atomicXor(&bits[x >> 5], m);
        if (old == 0u || old == m) atomicXor(&summ[(x >> 5) >> 5], 1u << ((x >> 5) & 31));
    }
    const unsigned long long mok = __ballot(ok);
    delta += 2 * __popcll(__ballot(ok && !(old & m))) - __popcll(mok);
    if (mok) lmax = max(lmax, __builtin_amdgcn_readlane(x, 63 - __clzll((long long)mok)));
    return !__ballot(need && !ok);
}

// Window words of the pipelined reducer's window form: the arena beside tw
// mark words, a multiple of 1024 (whole summary words), at most cap units
// of 1024 words (0: no cap).
__host__ __device__ __forceinline__ int pipe_window_words(int arena, int tw, int cap) {
    int ww = ((arena - tw) * 32 / 33) & ~1023;
    if (cap > 0 && ww > cap * 1024) ww = cap * 1024;
    return ww;
}

// ---- the pipelined reducer's window form ------------------------------
//
// For segments whose bitmap does not fit a wavefront's LDS arena: as the
// serial reducer's window (Win, above), by one wavefront: the rows of ww
// words from the column's first pivot on live in LDS with their summary,
// rows past them are flipped in the wavefront's global slot by atomics not
// waited for, their words marked in LDS; the window moves to the next
// marked nonzero word when it empties; the length is an upper bound until
// the column ends.  A wavefront's LDS operations take effect in issue
// order, so no barrier is needed; its global flips are drained
// (s_waitcnt vmcnt(0)) before the slot is read.

// Flips rows src[b0 + u * 64 + lane] (u < kPipeUnroll, < n), all loads
// first; hi = max(hi, rows flipped); false when a row lies outside
// [32 wlo, rows).
__device__ __forceinline__ bool pw_flip_chunk(const Win& W, const unsigned* src, int n, int b0,
                                              long long row_lo, int rows, int& hi) {
    const int lane = fph_lane();
    int x[kPipeUnroll];
    bool ok[kPipeUnroll];
    bool bad = false;
#pragma unroll
    for (int u = 0; u < kPipeUnroll; u++) {
        const int e = b0 + u * FPH_WAVE + lane;
        ok[u] = false;
        x[u] = 0;
        if (e < n) {
            const long long xx = (long long)ld_coh(src + e) - row_lo;
            ok[u] = xx >= 32LL * W.wlo && xx < rows;
            bad |= !ok[u];
            x[u] = ok[u] ? (int)xx : 0;
        }
    }
#pragma unroll
    for (int u = 0; u < kPipeUnroll; u++)
        if (ok[u]) {
            win_flip(W, x[u]);
            hi = max(hi, x[u]);
        }
    return !__ballot(bad);
}

// Lowest set row of the window at or above window word j0 (-1 none, -2 the
// summary and the bitmap disagree): win_search by one wavefront.

// The first nonzero marked global word at or past g0 and at most gmax, or
// -1; marks of zero words passed are cleared.  After vm_drain().
__device__ int pw_next_word(Win& W, int g0, int gmax) {
    const int lane = fph_lane();
    const int te = min(W.tw - 1, gmax >> 5);
    for (int tb = g0 >> 5; tb <= te; tb += FPH_WAVE) {
        const int tw = tb + lane;
        unsigned bits = tw <= te ? W.ts[tw] : 0u;
        if (tw == (g0 >> 5)) bits &= ~0u << (g0 & 31);
        int first = 0x7FFFFFFF;
        unsigned zero = 0u;
        for (unsigned m = bits; m; m &= m - 1) {
            const int g = tw * 32 + __builtin_ctz(m);
            if (ld_agent(&W.gbits[g])) {
                first = g;
                break;
            }
            zero |= m & (~m + 1u);
        }
        if (zero) W.ts[tw] &= ~zero;
        for (int o = FPH_WAVE / 2; o; o >>= 1) first = min(first, __shfl_xor(first, o));
        if (first != 0x7FFFFFFF) return first;
    }
    return -1;
}

// Moves the (empty) window to global word gn: the marked words of [gn, gn +
// ww) copied in, zeroed and unmarked.
__device__ void pw_move(Win& W, int gn) {
    const int lane = fph_lane();
    W.wlo = gn;
    const int ge = min(gn + W.ww, W.gw);
    for (int g = gn + lane; g < ge; g += FPH_WAVE) {
        if ((W.ts[g >> 5] >> (g & 31)) & 1u) {
            const unsigned v = ld_agent(&W.gbits[g]);
            if (v) {
                const int j = g - gn;
                W.wb[j] = v;
                atomicOr(&W.wsum[j >> 5], 1u << (j & 31));
                __hip_atomic_store(&W.gbits[g], 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    for (int tw = (gn >> 5) + lane; tw <= (ge - 1) >> 5; tw += FPH_WAVE) {
        const int lo = max(gn - tw * 32, 0), hi = min(ge - tw * 32, 32);  // bits [lo, hi)
        const unsigned m = (hi == 32 ? ~0u : (1u << hi) - 1u) & (~0u << lo);
        atomicAnd(&W.ts[tw], ~m);
    }
    vm_drain();
}

// Set rows of window summary words [s0, s1] and marked global words of mark
// words [m0, m1].  After vm_drain().
__device__ int pw_count(const Win& W, int s0, int s1, int m0, int m1) {
    const int lane = fph_lane();
    int cnt = 0;
    for (int sw = s0 + lane; sw <= s1; sw += FPH_WAVE)
        for (unsigned m = W.wsum[sw]; m; m &= m - 1) cnt += __popc(W.wb[sw * 32 + __builtin_ctz(m)]);
    for (int tw = m0 + lane; tw <= m1; tw += FPH_WAVE) {
        unsigned m = W.ts[tw];
        while (m) {
            int g[8];
            unsigned v[8];
            const int k = win_load8(W, tw, m, g, v);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < k) cnt += __popc(v[q]);
        }
    }
    int tot;
    (void)wave_excl_scan(cnt, &tot);
    return tot;
}

// Writes the column's rows ascending to dst (unless null) and clears it:
// window summary words [s0, s1], marked global words of mark words [m0, m1]
// (zeroed, unmarked).  Returns the rows found.  After vm_drain().
__device__ int pw_drain(Win& W, int s0, int s1, int m0, int m1, unsigned* dst,
                        long long row_lo) {
    const int lane = fph_lane();
    int written = 0;
    for (int sb = s0; sb <= s1; sb += FPH_WAVE) {
        const int sw = sb + lane;
        const unsigned val = sw <= s1 ? W.wsum[sw] : 0u;
        if (__ballot(val != 0u) == 0ull) continue;
        int cnt = 0;
        for (unsigned m = val; m; m &= m - 1) cnt += __popc(W.wb[sw * 32 + __builtin_ctz(m)]);
        int tot;
        int at = written + wave_excl_scan(cnt, &tot);
        for (unsigned m = val; m; m &= m - 1) {
            const int j = sw * 32 + __builtin_ctz(m);
            unsigned bw = W.wb[j];
            if (dst)
                for (; bw; bw &= bw - 1)
                    st_coh(dst + at++,
                           (unsigned)(row_lo + (long long)(W.wlo + j) * 32 + __builtin_ctz(bw)));
            W.wb[j] = 0u;
        }
        if (val) W.wsum[sw] = 0u;
        written += tot;
    }
    for (int tb = m0; tb <= m1; tb += FPH_WAVE) {
        const int tw = tb + lane;
        const unsigned bits = tw <= m1 ? W.ts[tw] : 0u;
        if (__ballot(bits != 0u) == 0ull) continue;
        int cnt = 0;
        for (unsigned m = bits; m;) {
            int g[8];
            unsigned v[8];
            const int k = win_load8(W, tw, m, g, v);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < k) cnt += __popc(v[q]);
        }
        int tot;
        int at = written + wave_excl_scan(cnt, &tot);
        for (unsigned m = bits; m;) {
            int g[8];
            unsigned v[8];
            const int k = win_load8(W, tw, m, g, v);
            for (int q = 0; q < k; q++) {
                if (dst)
                    for (unsigned bw = v[q]; bw; bw &= bw - 1)
                        st_coh(dst + at++, (unsigned)(row_lo + (long long)g[q] * 32 + __builtin_ctz(bw)));
                __hip_atomic_store(&W.gbits[g[q]], 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (bits) W.ts[tw] = 0u;
        written += tot;
    }
    vm_drain();
    return written;
}

__device__ __forceinline__ int wave_max(int v) {
    for (int o = FPH_WAVE / 2; o; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// pipe_column's window form: column order[r] of segment s (then any
// candidate a settling run hands back), the window and marks in smem (all
// zero on entry and on return), the slot gslot.
__device__ bool pipe_column_win(const PipeArgs& P, int s, long long r, unsigned* smem,
                                unsigned* gslot, Alloc& al, int& nxor, long long& merged,
                                int& longest) {
    const RedArgs& a = P.a;
    const int lane = fph_lane();
    const long long* sp = P.seg_pool + kSegPoolWords * s;
    const long long row_lo = sp[2];
    const int rows = (int)sp[3];
    unsigned long long* own = a.own + (row_lo - a.row_base);  // by local row
    Win W;
    W.gw = (rows + 31) / 32;
    W.tw = (W.gw + 31) / 32;
    W.ww = pipe_window_words(P.arena, W.tw, P.win_cap);
    W.ts = smem;
    W.wsum = smem + W.tw;
    W.wb = smem + W.tw + (W.ww >> 5);
    W.gbits = gslot;
    long long pos = r;
    bool adopted = false;
    while (true) {
        const int c = (int)P.order[pos];
        const unsigned* src;
        int n0;
        if (!adopted) {
            const int lo_c = a.off[c];
            src = a.data + lo_c;
            n0 = min(a.off[c + 1] - lo_c, P.maxcol);
        } else {
            const unsigned sg0 = ld_coh(&P.stg[pos]);
            src = a.pool + sg0 + 1;
            n0 = (int)ld_coh(a.pool + sg0);
        }
        int len = 0, p = 0;
        if (n0 > 0) {
            const long long p0 = (long long)ld_coh(src) - row_lo;
            if (p0 < 0 || p0 >= rows) return pipe_fail(P, 3);
            p = (int)p0;
            W.wlo = p >> 5;
            int hi = p;
            bool bad = false;
            for (int b0 = 0; b0 < n0; b0 += kPipeUnroll * FPH_WAVE)
                if (!pw_flip_chunk(W, src, n0, b0, row_lo, rows, hi)) bad = true;
            long long ub = n0;  // >= the column's length
            bool zero = false;
            while (true) {
                const unsigned long long v = ld_coh64(&own[p]);
                if (v == kNoOwner) break;
                // [len, rows...]: the length and the first 63 rows in one round trip
                const unsigned* ob = a.pool + (v & 0xFFFFFFFFull);
                const unsigned w0 = ld_coh(ob + lane);
                const int nb = __builtin_amdgcn_readlane((int)w0, 0);
                {
                    const bool need = lane >= 1 && lane <= nb;
                    const long long xx = (long long)w0 - row_lo;
                    const bool ok = need && xx >= 32LL * W.wlo && xx < rows;
                    if (need && !ok) bad = true;
                    if (ok) {
                        win_flip(W, (int)xx);
                        hi = max(hi, (int)xx);
                    }
                }
                for (int b0 = FPH_WAVE - 1; b0 < nb; b0 += kPipeUnroll * FPH_WAVE)
                    if (!pw_flip_chunk(W, ob + 1, nb, b0, row_lo, rows, hi)) bad = true;
                merged += nb;
                nxor++;
                ub += nb;
                if (ub > P.maxcol) {  // count the rows
                    vm_drain();
                    const int hm = wave_max(hi);
                    const int n = pw_count(W, 0, (W.ww >> 5) - 1, (W.wlo + W.ww) >> 5,
                                           min(W.tw - 1, hm >> 10));
                    if (n > P.maxcol) return pipe_fail(P, 4);  // the reference truncates
                    ub = n;
                }
                int np = win_search(W, (p >> 5) - W.wlo);
                if (np == -1) {  // the window is empty: move it to the next set row
                    vm_drain();
                    const int gn = pw_next_word(W, W.wlo + W.ww, wave_max(hi) >> 5);
                    if (gn < 0) {
                        zero = true;
                        break;
                    }
                    pw_move(W, gn);
                    const unsigned w0n = W.wb[0];
                    np = w0n ? gn * 32 + __builtin_ctz(w0n) : -1;
                }
                if (!(np > p)) return pipe_fail(P, 3);
                p = np;
            }
            if (__ballot(bad)) return pipe_fail(P, 3);
            vm_drain();
            if (!zero) {
                const int hm = wave_max(hi);
                const int s0 = ((p >> 5) - W.wlo) >> 5;
                const int s1 = min((W.ww >> 5) - 1, ((hm >> 5) - W.wlo) >> 5);
                const int m0 = (W.wlo + W.ww) >> 5, m1 = min(W.tw - 1, hm >> 10);
                len = pw_count(W, s0, s1, m0, m1);
                longest = max(longest, len);
                const long long at = pool_alloc(a, al, len + 1);
                if (at < 0) return pipe_fail(P, 1);
                const int written = pw_drain(W, s0, s1, m0, m1, a.pool + at + 1, row_lo);
                if (written != len) return pipe_fail(P, 3);
                if (lane == 0) {
                    st_coh(a.pool + at, (unsigned)len);
                    st_coh(&P.stg[pos], (unsigned)at);
                }
            }
        }
        // publish (rows first), then settle from here as pipe_column does
        vm_drain();
        if (lane == 0)
            st_coh64(&P.fin[pos], ((unsigned long long)(unsigned)p << 32) | (unsigned)len);
        vm_drain();
        if (ld_coh_i(&P.ctl[kPipeAbort])) return false;
        const unsigned long long prev = pipe_prev(P, s, pos);
        if (adopted) {
            if (prev == kNoRecord) return pipe_fail(P, 5);
        } else {
            if (prev == kNoRecord) return true;
            int won = 0;
            if (lane == 0) won = atomicCAS(&P.claim[pos], 0, 2) == 0;
            if (!__shfl(won, 0)) return true;
        }
        const long long u = pipe_hold(P, s, pos, prev);
        if (u < 0) return true;
        pos = u;
        adopted = true;
    }
}

// Column order[r] of segment s, by one wavefront; bits/summ as in
// serial_bitmap (all zero on entry and on return).  False on error or abort.
template <bool G>
__device__ bool pipe_column(const PipeArgs& P, int s, long long r, unsigned* bits,
                            unsigned* summ, Alloc& al, int& nxor, long long& merged,
                            int& longest) {
    const RedArgs& a = P.a;
    const int lane = fph_lane();
    const long long* sp = P.seg_pool + kSegPoolWords * s;
    const long long row_lo = sp[2];
    const int rows = (int)sp[3];
    const int S = ((rows + 31) / 32 + 31) / 32;
    unsigned long long* own = a.own + (row_lo - a.row_base);  // by local row
    // position r's column from the CSR; then, while a settling run hands one
    // back, a withdrawn candidate from its staged rows
    long long pos = r;
    bool adopted = false;
    while (true) {
        const int c = (int)P.order[pos];
        const unsigned* src;
        int n0;
        if (!adopted) {
            const int lo_c = a.off[c];
            src = a.data + lo_c;
            n0 = min(a.off[c + 1] - lo_c, P.maxcol);
        } else {
            const unsigned sg0 = ld_coh(&P.stg[pos]);
            src = a.pool + sg0 + 1;
            n0 = (int)ld_coh(a.pool + sg0);
        }
        int len = 0, p = 0, phigh = 0;
        if (n0 > 0) {
            int d0 = 0, lm0 = 0;
            bool bad = false;
            for (int b0 = 0; b0 < n0; b0 += kPipeUnroll * FPH_WAVE)
                if (!pipe_flip_chunk<G>(src, n0, b0, row_lo, rows, bits, summ, d0, lm0))
                    bad = true;
            if (bad) return pipe_fail(P, 3);
            len = n0;
            p = (int)((long long)ld_coh(src) - row_lo);
            phigh = (int)((long long)ld_coh(src + n0 - 1) - row_lo);
        }
        // add stored columns while the pivot has one
        while (len > 0) {
            unsigned long long v = ld_coh64(&own[p]);
            if (v == kNoOwner) break;
            // [len, rows...]: the length and the first 63 rows in one round trip
            // (a read past the column's end is a pool word, ignored; the pool
            // has 64 words of slack past its capacity)
            const unsigned* ob = a.pool + (v & 0xFFFFFFFFull);
            const unsigned w0 = ld_coh(ob + lane);
            const int nb = __builtin_amdgcn_readlane((int)w0, 0);
            int delta = 0, lmax = 0;
            bool badw = !pipe_flip_first<G>(w0, nb, row_lo, rows, bits, summ, delta, lmax);
            for (int b0 = FPH_WAVE - 1; b0 < nb; b0 += kPipeUnroll * FPH_WAVE) {
                if (!pipe_flip_chunk<G>(ob + 1, nb, b0, row_lo, rows, bits, summ, delta, lmax))
                    badw = true;
            }
            if (badw) return pipe_fail(P, 3);
            merged += 2LL * len + nb + delta;
            len += delta;
            phigh = max(phigh, lmax);
            nxor++;
            if (len > P.maxcol) return pipe_fail(P, 4);  // the reference truncates: list form
            if (len == 0) break;
            // the next pivot: the lowest set bit above p, through the summary
            const int ws = p >> 5;
            const int s_lo = ws >> 5, s_hi = min(phigh >> 10, S - 1);
            int np = -1;
            for (int i0 = s_lo; i0 <= s_hi; i0 += FPH_WAVE) {
                const int i = i0 + lane;
                unsigned val = i <= s_hi ? summ[i] : 0u;
                if (i == s_lo) val &= ~0u << (ws & 31);
                const unsigned long long hit = __ballot(val != 0u);
                if (hit) {
                    const int l = __ffsll((long long)hit) - 1;
                    int cand = -1;
                    if (lane == l) {
                        const int w = i * 32 + __builtin_ctz(val);
                        const unsigned bw = bm_load<G>(&bits[w]);
                        cand = bw ? w * 32 + __builtin_ctz(bw) : -1;
                    }
                    np = __builtin_amdgcn_readlane(cand, l);
                    break;
                }
            }
            if (!(np > p)) return pipe_fail(P, 3);
            p = np;
        }
        // the candidate's rows staged in ascending order, the bitmap cleared
        long long at = 0;
        if (len > 0) {
            longest = max(longest, len);
            at = pool_alloc(a, al, len + 1);
            if (at < 0) return pipe_fail(P, 1);
            int written = 0;
            const int sw_hi = min(phigh >> 10, S - 1);
            for (int sw0 = p >> 10; sw0 <= sw_hi; sw0 += FPH_WAVE) {
                const int sw = sw0 + lane;
                const unsigned val = sw <= sw_hi ? summ[sw] : 0u;
                if (__ballot(val != 0u) == 0ull) continue;  // 65,536 rows, none set
                // a lane's words 8 at a time, their loads in flight together;
                // the first 8 are kept for the writing pass
                int cnt = 0;
                unsigned c8[8];
                int cw8[8];
                unsigned m = val;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    cw8[q] = m ? sw * 32 + __builtin_ctz(m) : -1;
                    c8[q] = m ? bm_load<G>(&bits[cw8[q]]) : 0u;
                    m &= m - 1;
                }
                const unsigned rest = m;  // words past the first 8
#pragma unroll
                for (int q = 0; q < 8; q++) cnt += __popc(c8[q]);
                while (m) {
                    unsigned v8[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        v8[q] = m ? bm_load<G>(&bits[sw * 32 + __builtin_ctz(m)]) : 0u;
                        m &= m - 1;
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) cnt += __popc(v8[q]);
                }
                int tot;
                long long o = wave_excl_scan(cnt, &tot) + written;
                for (int q = 0; q < 8; q++) {
                    if (cw8[q] < 0) break;
                    for (unsigned bw = c8[q]; bw; bw &= bw - 1)
                        st_coh(a.pool + at + 1 + o++,
                               (unsigned)(row_lo + cw8[q] * 32 + __builtin_ctz(bw)));
                    bm_zero<G>(&bits[cw8[q]]);
                }
                for (m = rest; m;) {
                    unsigned v8[8];
                    int w8[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        w8[q] = m ? sw * 32 + __builtin_ctz(m) : -1;
                        v8[q] = m ? bm_load<G>(&bits[w8[q]]) : 0u;
                        m &= m - 1;
                    }
                    for (int q = 0; q < 8; q++) {
                        if (w8[q] < 0) break;
                        for (unsigned bw = v8[q]; bw; bw &= bw - 1)
                            st_coh(a.pool + at + 1 + o++,
                                   (unsigned)(row_lo + w8[q] * 32 + __builtin_ctz(bw)));
                        bm_zero<G>(&bits[w8[q]]);
                    }
                }
                if (val) summ[sw] = 0u;
                written += tot;
            }
            if (written != len) return pipe_fail(P, 3);
            if (lane == 0) {
                st_coh(a.pool + at, (unsigned)len);
                st_coh(&P.stg[pos], (unsigned)at);
            }
        }
        // publish (rows first), then settle from here if the position before
        // has settled: this wavefront or the run that reached it, not both
        vm_drain();
        if (lane == 0)
            st_coh64(&P.fin[pos], ((unsigned long long)(unsigned)p << 32) | (unsigned)len);
        vm_drain();
        if (ld_coh_i(&P.ctl[kPipeAbort])) return false;
        const unsigned long long prev = pipe_prev(P, s, pos);
        if (adopted) {
            // claimed when it was withdrawn, after the record before it
            if (prev == kNoRecord) return pipe_fail(P, 5);
        } else {
            if (prev == kNoRecord) return true;
            int won = 0;
            if (lane == 0) won = atomicCAS(&P.claim[pos], 0, 2) == 0;
            if (!__shfl(won, 0)) return true;
        }
        const long long u = pipe_hold(P, s, pos, prev);
        if (u < 0) return true;
        pos = u;  // a withdrawn candidate: add what was stored for its pivot
        adopted = true;
    }
}

// One wavefront per workgroup.  Each XCD keeps up to kPipeSlots segments
// open (an XCD opens whole segments, most residual columns first; all the
// columns of a segment are taken by wavefronts of its XCD).  A wavefront
// takes the next column of the first open segment of its XCD with columns
// left and fewer than `inflight` of them unsettled; when none qualifies it
// opens the next segment in a free slot (or one whose columns are all
// handed out).  The cap holds throughout: a column handed out far ahead of
// the settled positions is likely to stop at a pivot stored later and be
// taken over again (the 256-cloud batch: 92 ms with a cap of 32, 400 ms
// with none once every segment is open).
__global__ __launch_bounds__(FPH_WAVE) void k_reduce_pipe(PipeArgs P) {
    extern __shared__ unsigned smem[];
    const int lane = fph_lane();
    const int xcc = (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 7);
    for (int i = lane; i < P.arena; i += FPH_WAVE) smem[i] = 0u;
    wave_sync();
    unsigned* gslot = P.gbits ? P.gbits + (long long)blockIdx.x * P.wmax : nullptr;
    int nxor = 0, longest = 0;
    long long merged = 0;
    Alloc al = make_alloc(P.chunk);
    Deadline idle;  // a wavefront finding no work for spin_limit: status 5 (a scheduling bug)
    int* slots = P.ctl + xcc * kPipeSlots;
    while (!ld_coh_i(&P.ctl[kPipeAbort])) {
        const bool all_open = ld_coh_i(&P.ctl[kPipeNext]) >= P.nseg;
        int seg = -1, pos = -1, freeable = -1, freeable_val = -1;
        bool left = false;  // an open segment of this XCD has columns left
        for (int k = 0; k < kPipeSlots && seg < 0; k++) {
            const int cur = ld_coh_i(slots + k);
            if (cur == -1) {
                if (freeable < 0) freeable = k, freeable_val = cur;
                continue;
            }
            if (cur < 0) {  // being opened
                left = true;
                continue;
            }
            const int lo = P.seg_lo[cur], R = P.seg_hi[cur] - lo;
            const int nx = ld_coh_i(&P.seg_next[cur]);
            if (nx >= R) {
                if (freeable < 0) freeable = k, freeable_val = cur;
                continue;
            }
            left = true;
            const int rw = (int)((P.seg_pool[kSegPoolWords * cur + 3] + 31) / 32);
            const int cap = rw + (rw + 31) / 32 <= P.arena && !P.force_global ? P.inflight_lds
                                                                              : P.inflight;
            if (nx - ld_coh_i(&P.seg_done[cur]) >= cap) continue;
            int i = 0;
            if (lane == 0) i = atomicAdd(&P.seg_next[cur], 1);
            i = __shfl(i, 0);
            if (i < R) {
                seg = cur;
                pos = lo + i;
            }
        }
        if (seg >= 0) {
            idle.t0 = -1;
            const int rows = (int)P.seg_pool[kSegPoolWords * seg + 3];
            const int W = (rows + 31) / 32, S = (W + 31) / 32;
            bool ok;
            if (W + S <= P.arena && !P.force_global)
                ok = pipe_column<false>(P, seg, pos, smem + S, smem, al, nxor, merged, longest);
            else if (P.window && pipe_window_words(P.arena, S, P.win_cap) >= kWinMinWords)
                ok = pipe_column_win(P, seg, pos, smem, gslot, al, nxor, merged, longest);
            else
                ok = pipe_column<true>(P, seg, pos, gslot, smem, al, nxor, merged, longest);
            if (!ok) break;
            continue;
        }
        if (all_open) {
            if (!left) break;  // nothing left for this XCD
            if (idle.expired(P.spin_limit)) {
                pipe_fail(P, 5);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        // open the next segment in a free slot (only if the slot still holds
        // what the scan found there: -1 or a segment with every column taken)
        if (freeable >= 0 && lane == 0) {
            if (atomicCAS(slots + freeable, freeable_val, -2) == freeable_val) {
                const int t = atomicAdd(&P.ctl[kPipeNext], 1);
                __hip_atomic_store(slots + freeable, t < P.nseg ? P.seg_order[t] : -1,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (freeable < 0) {
            if (idle.expired(P.spin_limit)) {
                pipe_fail(P, 5);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        wave_sync();
    }
    if (lane == 0) {
        if (nxor) atomicAdd(&P.a.status[2], nxor);
        if (merged) count_merged(P.a, merged);
        if (longest) atomicMax(&P.a.status[3], longest);
    }
}

// Positions of each segment in order (contiguous: order is descending and a
// segment is a column range), its cost rank, and the control words.
__global__ void k_pipe_prep(const unsigned* __restrict__ order, const int* __restrict__ res_count,
                            const int* __restrict__ seg_cols, int nseg, long long K,
                            int* __restrict__ seg_lo, int* __restrict__ seg_hi,
                            int* __restrict__ seg_next, int* __restrict__ seg_done,
                            int* __restrict__ ctl) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < kPipeCtl) ctl[s] = s < kPipeNext ? -1 : 0;
    if (s >= nseg) return;
    const int R = *res_count;
    const int c0 = seg_cols ? seg_cols[s] : 0;
    const int c1 = seg_cols ? seg_cols[s + 1] : (int)K;
    int lo = 0, hi = R;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)order[mid] >= c1) lo = mid + 1;
        else hi = mid;
    }
    int lo2 = lo, hi2 = R;
    while (lo2 < hi2) {
        const int mid = (lo2 + hi2) >> 1;
        if ((int)order[mid] >= c0) lo2 = mid + 1;
        else hi2 = mid;
    }
    seg_lo[s] = lo;
    seg_hi[s] = lo2;
    seg_next[s] = 0;
    seg_done[s] = 0;
}

// seg_order: segments by descending residual count (ties by index); up to
// kPipeRankMax segments, beyond that in index order.
constexpr int kPipeRankMax = 8192;

// Every position settled (a check of the protocol): status 5 otherwise.
__global__ void k_pipe_check(const int* __restrict__ seg_lo, const int* __restrict__ seg_hi,
                             const int* __restrict__ seg_done, int nseg, int* __restrict__ status) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nseg && seg_done[s] != seg_hi[s] - seg_lo[s]) atomicCAS(&status[0], 0, 5);
}

__global__ void k_pipe_rank(const int* __restrict__ seg_lo, const int* __restrict__ seg_hi,
                            int nseg, int* __restrict__ seg_order) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    if (nseg > kPipeRankMax) {
        seg_order[s] = s;
        return;
    }
    const int rs = seg_hi[s] - seg_lo[s];
    int rank = 0;
    for (int t = 0; t < nseg; t++) {
        const int rt = seg_hi[t] - seg_lo[t];
        rank += rt > rs || (rt == rs && t < s);
    }
    seg_order[rank] = s;
}

// The prefill of each segment (one workgroup each) and its record.
__global__ __launch_bounds__(kSerThreads) void k_pipe_prefill(RedArgs a, const int* seg_cols,
                                                             const long long* seg_pool, int exact,
                                                             unsigned long long* seg_state) {
    __shared__ SerShared sh;
    const int seg = blockIdx.x;
    const int c0 = seg_cols ? seg_cols[seg] : 0;
    const int c1 = seg_cols ? seg_cols[seg + 1] : (int)a.K;
    const long long used = serial_prefill(a, sh, c0, c1, seg_pool[kSegPoolWords * seg + 1]);
    if (threadIdx.x == 0) seg_state[seg] = exact ? 0ull : (unsigned long long)used << 32;
}

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

size_t sort_bytes(int64_t K) {
    size_t tmp = 0;
    (void)rocprim::radix_sort_keys_desc(nullptr, tmp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                        (unsigned)std::max<int64_t>(K, 1), 0, 32, (hipStream_t)0);
    return align_up(tmp);
}

}  // namespace

// order[0..res_count) = the residual columns in descending index.
int launch_order(const unsigned char* kind, const int* res_list, const int* res_count, int64_t K,
                 unsigned* keys, unsigned* order, void* sort_tmp, size_t sort_sz,
                 hipStream_t s) {
    if (K <= kOrderCompactMax) {
        k_order_desc<<<1, kOrdThreads, 0, s>>>(kind, K, order);
        FPH_LAUNCH_CHECK();
        return FPH_OK;
    }
    k_order_keys<<<fph_blocks(K, 256), 256, 0, s>>>(res_list, res_count, K, keys);
    FPH_LAUNCH_CHECK();
    FPH_CHECK_HIP(rocprim::radix_sort_keys_desc(sort_tmp, sort_sz, keys, order, (unsigned)K, 0,
                                                32, s));
    return FPH_OK;
}

extern "C" size_t fph_reduce_scratch_bytes(int64_t K) {
    const size_t k4 = align_up(4 * (size_t)std::max<int64_t>(K, 1));
    // desc, two worklists (the displaced ring), sort keys in/out, sort
    // temp, pool counter
    return 5 * k4 + sort_bytes(K) + 256;
}


extern "C" int fph_reduce_serial(int64_t K, const int32_t* col_off, const uint32_t* col_data,
                                 int64_t row_base, int64_t nrows, const int32_t* seg_cols,
                                 int64_t nseg, const int64_t* seg_pool, int64_t pool_start,
                                 const uint8_t* kind, uint64_t* own, int32_t* pivot,
                                 const int32_t* res_list, const int32_t* res_count,
                                 uint32_t* pool, int64_t pool_cap, void* scratch,
                                 size_t scratch_bytes, int32_t* status, int variant,
                                 int64_t max_col, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || nrows < 0 || pool_cap < 0 || !scratch || !status || nseg < 1 || !seg_pool)
        return FPH_ERR_ARG;
    if (max_col == 0) max_col = kSerialMaxCol;
    if (max_col < 1 || max_col > 0x7FFFFFFFLL) return FPH_ERR_ARG;
    if (nseg > 1 && !seg_cols) return FPH_ERR_ARG;
    if (scratch_bytes < fph_reduce_scratch_bytes(K)) return FPH_ERR_ARG;
    if (pool_start < 0 || pool_start > pool_cap || pool_cap > 0xFFFFFFFFLL) return FPH_ERR_ARG;
    const size_t k4 = align_up(4 * (size_t)std::max<int64_t>(K, 1));
    char* w = (char*)scratch;
    RedArgs a;
    a.off = col_off;
    a.data = col_data;
    a.row_base = row_base;
    a.nrows = nrows;
    a.K = K;
    a.kind = kind;
    a.own = (unsigned long long*)own;
    a.pivot = pivot;
    a.desc = (unsigned*)w;
    unsigned* keys = (unsigned*)(w + 3 * k4);
    unsigned* order = (unsigned*)(w + 4 * k4);
    void* sort_tmp = w + 5 * k4;
    size_t sort_sz = sort_bytes(K);
    a.pool_ptr = (unsigned long long*)(w + 5 * k4 + sort_sz);
    a.pool = pool;
    a.pool_cap = (unsigned long long)pool_cap;
    a.status = status;
    a.use_lines = 0;
    a.lines = nullptr;
    k_reduce_init<<<1, 64, 0, s>>>(status, a.pool_ptr, col_off, K, a);
    FPH_LAUNCH_CHECK();
    k_set_u64<<<1, 64, 0, s>>>(a.pool_ptr, (unsigned long long)pool_start);
    FPH_LAUNCH_CHECK();
    if (K > 0) {
        k_publish_apparent_flat<<<(unsigned)std::min<int64_t>(16384, (K + 255) / 256), 256, 0,
                                  s>>>(a, K);
        FPH_LAUNCH_CHECK();
        const int orc = launch_order(kind, res_list, res_count, K, keys, order, sort_tmp, sort_sz, s);
        if (orc) return orc;
        FPH_CHECK_HIP(hipFuncSetAttribute((const void*)k_reduce_serial,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)kSerLds));
        k_reduce_serial<<<(unsigned)nseg, kSerThreads, kSerLds, s>>>(
            a, order, res_count, seg_cols, (const long long*)seg_pool, variant, (int)max_col);
        FPH_LAUNCH_CHECK();
    }
    k_reduce_fini<<<1, 64, 0, s>>>(status, a.pool_ptr);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_reduce(int64_t K, const int32_t* col_off, const uint32_t* col_data,
                          int64_t row_base, int64_t nrows, const int32_t* seg_cols,
                          int64_t nseg, const uint8_t* kind, uint64_t* own, int32_t* pivot,
                          const int32_t* res_list, const int32_t* res_count, uint32_t* pool,
                          int64_t pool_cap, void* scratch, size_t scratch_bytes,
                          int32_t* status, int wait_seconds, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || nrows < 0 || pool_cap < 0 || !scratch || !status || nseg < 1) return FPH_ERR_ARG;
    if (nseg > 1 && !seg_cols) return FPH_ERR_ARG;
    if (scratch_bytes < fph_reduce_scratch_bytes(K)) return FPH_ERR_ARG;
    const size_t k4 = align_up(4 * (size_t)std::max<int64_t>(K, 1));
    char* w = (char*)scratch;
    RedArgs a;
    a.off = col_off;
    a.data = col_data;
    a.row_base = row_base;
    a.nrows = nrows;
    a.K = K;
    a.kind = kind;
    a.own = (unsigned long long*)own;
    a.pivot = pivot;
    a.desc = (unsigned*)w;
    int* wl0 = (int*)(w + k4);
    unsigned* keys = (unsigned*)(w + 3 * k4);
    unsigned* order = (unsigned*)(w + 4 * k4);
    void* sort_tmp = w + 5 * k4;
    size_t sort_sz = sort_bytes(K);
    a.pool_ptr = (unsigned long long*)(w + 5 * k4 + sort_sz);
    a.pool = pool;
    a.pool_cap = (unsigned long long)pool_cap;
    a.status = status;
    // one segment: the queue may span several workgroups (autotuned); with
    // bitmap working columns it also takes phase A's additions (cheaper
    // there than as list merges), starts columns from the CSR and keeps
    // owner lines
    const int qwgs = nseg == 1 ? fph_tune_value(FPH_TUNE_QUEUE_WGS, K, 0) : 1;
    const int bitmap = (nrows + 31) / 32 <= kArenaB ? 1 : 0;
    const int from_csr = qwgs > 1 && bitmap;
    a.use_lines = from_csr;
    a.lines = nullptr;
    if (!(from_csr && K > 0)) {  // else k_queue_prep does it
        k_reduce_init<<<1, 64, 0, s>>>(status, a.pool_ptr, col_off, K, a);
        FPH_LAUNCH_CHECK();
    }
    if (K == 0) {
        k_reduce_fini<<<1, 64, 0, s>>>(status, a.pool_ptr);
        FPH_LAUNCH_CHECK();
        return FPH_OK;
    }
    // ring of displaced owners: K 64-bit words over the two worklist areas
    unsigned long long* ring = (unsigned long long*)wl0;
    int* qs = (int*)(w + 5 * k4 + sort_sz + 64);  // queue state, in the 256-byte tail
    if (from_csr) {  // before k_publish_apparent, which writes owner lines
        k_queue_prep<<<1024, 256, 0, s>>>(a, ring, qs, res_count, pool_cap + 64);
        FPH_LAUNCH_CHECK();
    }
    const unsigned grid_w =
        (unsigned)std::min<int64_t>(1024, (K + kPhaseAWaves - 1) / kPhaseAWaves);
    k_publish_apparent<<<grid_w, kPhaseAWaves * FPH_WAVE, 0, s>>>(a, K, 1);
    const size_t lds_a = kPhaseAWaves * kArenaA * 4;
    if (!from_csr)
        k_reduce_chain<<<grid_w, kPhaseAWaves * FPH_WAVE, lds_a, s>>>(a, res_list, res_count);
    FPH_LAUNCH_CHECK();
    const int orc = launch_order(kind, res_list, res_count, K, keys, order, sort_tmp, sort_sz, s);
    if (orc) return orc;
    if (!from_csr) {
        k_fill_u64<<<fph_blocks(K, 256), 256, 0, s>>>(ring, K, ~0ull);
        FPH_LAUNCH_CHECK();
    }
    const size_t lds_b = kPhaseBWaves * kArenaB * 4;
    FPH_CHECK_HIP(hipFuncSetAttribute((const void*)k_reduce_queue,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
    // no-progress bound of a waiting wavefront: wait_seconds (<= 0: 120 s)
    // of the device wall clock
    int dev = 0, khz = 0;
    FPH_CHECK_HIP(hipGetDevice(&dev));
    FPH_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const long long secs = wait_seconds > 0 ? (long long)wait_seconds : 120LL;
    const long long limit_ticks = secs * (long long)std::max(khz, 1) * 1000LL;
    if (qwgs > 1) {
        if (!from_csr) {
            k_fill_unwritten<<<1024, 256, 0, s>>>(pool, a.pool_ptr, pool_cap + 64);
            k_queue_state_init<<<1, 64, 0, s>>>(qs, res_count, status);
            FPH_LAUNCH_CHECK();
        }
        // 8 XCDs receive workgroups in turn: 8 x qwgs puts about qwgs on each;
        // bitmap working columns when the rows fit a wavefront's arena
        if (bitmap) {
            FPH_CHECK_HIP(hipFuncSetAttribute((const void*)k_reduce_queue_dev<true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds_b));
            k_reduce_queue_dev<true><<<(unsigned)(8 * qwgs), kPhaseBWaves * FPH_WAVE, lds_b, s>>>(
                a, order, res_count, ring, qs, limit_ticks, from_csr);
        } else {
            FPH_CHECK_HIP(hipFuncSetAttribute((const void*)k_reduce_queue_dev<false>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds_b));
            k_reduce_queue_dev<false><<<(unsigned)(8 * qwgs), kPhaseBWaves * FPH_WAVE, lds_b, s>>>(
                a, order, res_count, ring, qs, limit_ticks, from_csr);
        }
        FPH_LAUNCH_CHECK();
        if (!from_csr) k_final_reset<<<fph_blocks(K, 256), 256, 0, s>>>(a);
        // also k_reduce_fini's work
        k_final_own<<<fph_blocks(std::max<int64_t>(nrows, 1), 256), 256, 0, s>>>(a);
        FPH_LAUNCH_CHECK();
        return FPH_OK;
    }
    k_reduce_queue<<<(unsigned)nseg, kPhaseBWaves * FPH_WAVE, lds_b, s>>>(
        a, order, res_count, seg_cols, ring, limit_ticks);
    FPH_LAUNCH_CHECK();
    k_reduce_fini<<<1, 64, 0, s>>>(status, a.pool_ptr);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_reduce_replay(int64_t K, const int32_t* col_off, const uint32_t* col_data,
                                 int64_t row_base, int64_t nrows, const uint8_t* kind,
                                 const uint64_t* own, const int32_t* pivot,
                                 const int32_t* res_list, const int32_t* res_count,
                                 uint32_t* pool, int64_t pool_cap, void* scratch,
                                 size_t scratch_bytes, int32_t* status, int wait_seconds,
                                 void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || nrows < 0 || pool_cap < 0 || !scratch || !status || !pool || !res_count)
        return FPH_ERR_ARG;
    if (scratch_bytes < fph_reduce_scratch_bytes(K)) return FPH_ERR_ARG;
    // working columns as bitmaps in the wavefront's arena only
    if ((nrows + 31) / 32 > kArenaB) return FPH_ERR_LIMIT;
    if (pool_cap > 0xFFFFFFFFLL - 64) return FPH_ERR_LIMIT;
    const size_t k4 = align_up(4 * (size_t)std::max<int64_t>(K, 1));
    char* w = (char*)scratch;
    RedArgs a;
    a.off = col_off;
    a.data = col_data;
    a.row_base = row_base;
    a.nrows = nrows;
    a.K = K;
    a.kind = kind;
    a.own = const_cast<unsigned long long*>(reinterpret_cast<const unsigned long long*>(own));
    a.pivot = const_cast<int*>(pivot);
    a.desc = (unsigned*)w;
    unsigned* keys = (unsigned*)(w + 3 * k4);
    unsigned* order = (unsigned*)(w + 4 * k4);
    void* sort_tmp = w + 5 * k4;
    size_t sort_sz = sort_bytes(K);
    a.pool_ptr = (unsigned long long*)(w + 5 * k4 + sort_sz);
    int* qs = (int*)(w + 5 * k4 + sort_sz + 64);
    a.pool = pool;
    a.pool_cap = (unsigned long long)pool_cap;
    a.status = status;
    a.use_lines = 0;
    a.lines = nullptr;
    k_replay_init<<<1, 64, 0, s>>>(status, a.pool_ptr, qs);
    FPH_LAUNCH_CHECK();
    if (K == 0) return FPH_OK;
    k_fill_u32<<<fph_blocks(K, 256), 256, 0, s>>>(a.desc, K, kUnwritten);
    k_fill_unwritten<<<1024, 256, 0, s>>>(pool, a.pool_ptr, pool_cap + 64);
    k_replay_apparent<<<(unsigned)std::min<int64_t>(1024, fph_blocks(K, 256)), 256, 0, s>>>(
        kind, col_off, K, status);
    FPH_LAUNCH_CHECK();
    const int orc = launch_order(kind, res_list, res_count, K, keys, order, sort_tmp, sort_sz, s);
    if (orc) return orc;
    const size_t lds_b = kPhaseBWaves * kArenaB * 4;
    FPH_CHECK_HIP(hipFuncSetAttribute((const void*)k_replay,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
    int dev = 0, khz = 0;
    FPH_CHECK_HIP(hipGetDevice(&dev));
    FPH_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const long long secs = wait_seconds > 0 ? (long long)wait_seconds : 120LL;
    const long long limit_ticks = secs * (long long)std::max(khz, 1) * 1000LL;
    // 8 XCDs receive workgroups in turn: 8 x 16 puts 16 on the XCD taking part
    k_replay<<<8 * 16, kPhaseBWaves * FPH_WAVE, lds_b, s>>>(a, order, res_count, qs, limit_ticks);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

// ---- fph_reduce_pipe ---------------------------------------------------

namespace {

struct PipeLayout {
    int grid = 0, arena = 0;
    long long wmax = 0;
    size_t rec = 0, fin = 0, stg = 0, claim = 0, seg_lo = 0, seg_hi = 0, seg_next = 0,
           seg_done = 0, seg_order = 0, seg_state = 0, ctl = 0, gbits = 0, total = 0;
};

// Launch shape and workspace of fph_reduce_pipe.  A segment whose bitmap
// and summary fit kPipeLdsMax words keeps them in LDS; otherwise the summary
// is in LDS and the bitmap in a global slot per workgroup.  Returns false
// when a segment's summary alone does not fit.
bool pipe_layout(int64_t K, int64_t nseg, int64_t max_rows, int cus, bool force_global,
                 PipeLayout& L) {
    const long long W = (max_rows + 31) / 32, S = (W + 31) / 32;
    if (S > kPipeLdsMax) return false;
    // the arena holds the largest bitmap that fits, so that segments up to
    // that size use LDS; larger ones keep only their summary there
    const bool global = W + S > kPipeLdsMax || force_global;
    L.arena = (int)std::max<long long>(std::min<long long>(W + S, kPipeLdsMax), 64);
    // workgroups per CU by LDS (160 KB), at most 4; global slots at most 256 MiB
    int wpc = (int)std::min<long long>(4, std::max<long long>(1, 163840 / (4LL * L.arena + 1024)));
    if (global)
        while (wpc > 1 && (long long)cus * wpc * W * 4 > (256LL << 20)) wpc--;
    L.grid = cus * wpc;
    L.wmax = global ? W : 0;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t here = at;
        at += align_up(bytes);
        return here;
    };
    L.rec = take(8 * (size_t)std::max<int64_t>(K, 1));
    L.fin = take(8 * (size_t)std::max<int64_t>(K, 1));
    L.stg = take(4 * (size_t)std::max<int64_t>(K, 1));
    L.claim = take(4 * (size_t)std::max<int64_t>(K, 1));
    L.seg_lo = take(4 * (size_t)nseg);
    L.seg_hi = take(4 * (size_t)nseg);
    L.seg_next = take(4 * (size_t)nseg);
    L.seg_done = take(4 * (size_t)nseg);
    L.seg_order = take(4 * (size_t)nseg);
    L.seg_state = take(8 * (size_t)nseg);
    L.ctl = take(4 * kPipeCtl);
    L.gbits = take(global ? 4 * (size_t)L.grid * (size_t)W : 4);
    L.total = at;
    return true;
}

}  // namespace

extern "C" int fph_reduce_pipe(int64_t K, const int32_t* col_off, const uint32_t* col_data,
                               int64_t row_base, int64_t nrows, const int32_t* seg_cols,
                               int64_t nseg, const int64_t* seg_pool, int64_t max_rows,
                               int64_t pool_start, const uint8_t* kind, uint64_t* own,
                               int32_t* pivot, const int32_t* res_list, const int32_t* res_count,
                               uint32_t* pool, int64_t pool_cap, void* scratch,
                               size_t scratch_bytes, void* ws, size_t* ws_bytes, int32_t* status,
                               int variant, int64_t max_col, int wait_seconds, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (K < 0 || nrows < 0 || pool_cap < 0 || nseg < 1 || nseg > 0x7FFFFFFFLL || !ws_bytes ||
        max_rows < 0 || max_rows > 0x7FFFFFFFLL)
        return FPH_ERR_ARG;
    if (max_col == 0) max_col = kSerialMaxCol;
    if (max_col < 1 || max_col > 0x7FFFFFFFLL) return FPH_ERR_ARG;
    int dev = 0, cus = 0;
    FPH_CHECK_HIP(hipGetDevice(&dev));
    FPH_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PipeLayout L;
    if (!pipe_layout(K, nseg, max_rows, std::max(cus, 8), (variant & 4) != 0, L))
        return FPH_ERR_LIMIT;
    if (!ws) {
        *ws_bytes = L.total;
        return FPH_OK;
    }
    if (*ws_bytes < L.total || !scratch || !status || !seg_pool || !pool) return FPH_ERR_ARG;
    if (nseg > 1 && !seg_cols) return FPH_ERR_ARG;
    if (scratch_bytes < fph_reduce_scratch_bytes(K)) return FPH_ERR_ARG;
    if (pool_start < 0 || pool_start > pool_cap || pool_cap > 0xFFFFFFFFLL) return FPH_ERR_ARG;
    if (variant & 1) return FPH_ERR_ARG;  // the sorted-list form is fph_reduce_serial's
    const size_t k4 = align_up(4 * (size_t)std::max<int64_t>(K, 1));
    char* w = (char*)scratch;
    char* wb = (char*)ws;
    RedArgs a;
    a.off = col_off;
    a.data = col_data;
    a.row_base = row_base;
    a.nrows = nrows;
    a.K = K;
    a.kind = kind;
    a.own = (unsigned long long*)own;
    a.pivot = pivot;
    a.desc = (unsigned*)w;
    unsigned* keys = (unsigned*)(w + 3 * k4);
    unsigned* order = (unsigned*)(w + 4 * k4);
    void* sort_tmp = w + 5 * k4;
    size_t sort_sz = sort_bytes(K);
    a.pool_ptr = (unsigned long long*)(w + 5 * k4 + sort_sz);
    a.pool = pool;
    a.pool_cap = (unsigned long long)pool_cap;
    a.status = status;
    a.use_lines = 0;
    a.lines = nullptr;
    PipeArgs P;
    P.a = a;
    P.order = order;
    P.seg_pool = (const long long*)seg_pool;
    P.seg_lo = (int*)(wb + L.seg_lo);
    P.seg_hi = (int*)(wb + L.seg_hi);
    P.seg_next = (int*)(wb + L.seg_next);
    P.seg_done = (int*)(wb + L.seg_done);
    P.seg_order = (int*)(wb + L.seg_order);
    P.seg_state = (unsigned long long*)(wb + L.seg_state);
    P.rec = (unsigned long long*)(wb + L.rec);
    P.fin = (unsigned long long*)(wb + L.fin);
    P.stg = (unsigned*)(wb + L.stg);
    P.claim = (int*)(wb + L.claim);
    P.ctl = (int*)(wb + L.ctl);
    P.gbits = L.wmax > 0 ? (unsigned*)(wb + L.gbits) : nullptr;
    P.wmax = L.wmax;
    P.nseg = (int)nseg;
    P.arena = L.arena;
    P.maxcol = (variant & 2) ? 0x7FFFFFFF : (int)max_col;
    P.exact = (variant & 2) ? 1 : 0;
    P.force_global = (variant & 4) ? 1 : 0;
    P.window = (variant & 8) ? 0 : 1;
    P.win_cap = (variant >> 8) & 0xFF;
    // Columns of one segment in flight: more of them fill the XCD sooner but
    // stop more often at pivots whose owners settle later, to be taken over
    // at the settling front.  A quarter of an XCD's wavefronts for bitmaps in
    // LDS, an eighth where staging reads a global slot (a takeover costs more
    // there): at 512 wavefronts 48 and 16 (256 uneven clouds 92.5 -> 88 ms
    // against 32 for both; n=1000 dense cloud 14.2 -> 12.7 ms).  Bits 16..23
    // of variant: that many for both.
    P.inflight = std::max(8, L.grid / 32);
    P.inflight_lds = std::max(8, 3 * L.grid / 32);
    if ((variant >> 16) & 0xFF) P.inflight = P.inflight_lds = (variant >> 16) & 0xFF;
    // staging reservations: a quarter of the free pool spread over the
    // wavefronts, 256 to 65536 words each
    {
        const long long free_words = std::max<long long>(pool_cap - pool_start, 0);
        long long ch = free_words / (4LL * std::max(L.grid, 1));
        ch = std::min<long long>(std::max<long long>(ch, 256), 65536);
        P.chunk = (unsigned long long)(ch / kLine * kLine);
    }
    {
        int khz = 0;
        FPH_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        const long long secs = wait_seconds > 0 ? (long long)wait_seconds : 120LL;
        P.spin_limit = secs * (long long)std::max(khz, 1) * 1000LL;
    }
    k_reduce_init<<<1, 64, 0, s>>>(status, a.pool_ptr, col_off, K, a);
    FPH_LAUNCH_CHECK();
    k_set_u64<<<1, 64, 0, s>>>(a.pool_ptr, (unsigned long long)pool_start);
    FPH_LAUNCH_CHECK();
    if (K > 0) {
        k_publish_apparent_flat<<<(unsigned)std::min<int64_t>(16384, (K + 255) / 256), 256, 0,
                                  s>>>(a, K);
        FPH_LAUNCH_CHECK();
        const int orc = launch_order(kind, res_list, res_count, K, keys, order, sort_tmp, sort_sz, s);
        if (orc) return orc;
        k_fill_u64<<<fph_blocks(K, 256), 256, 0, s>>>(P.rec, K, kNoRecord);
        k_fill_u64<<<fph_blocks(K, 256), 256, 0, s>>>(P.fin, K, kNoRecord);
        k_fill_u32<<<fph_blocks(K, 256), 256, 0, s>>>((unsigned*)P.claim, K, 0u);
        if (P.gbits)
            k_fill_u32<<<fph_blocks((long long)L.grid * L.wmax, 256), 256, 0, s>>>(
                P.gbits, (long long)L.grid * L.wmax, 0u);
        k_pipe_prep<<<fph_blocks(std::max<int64_t>(nseg, kPipeCtl), 256), 256, 0, s>>>(
            order, res_count, seg_cols, (int)nseg, K, (int*)P.seg_lo, (int*)P.seg_hi, P.seg_next,
            P.seg_done, P.ctl);
        FPH_LAUNCH_CHECK();
        k_pipe_rank<<<fph_blocks(nseg, 256), 256, 0, s>>>(P.seg_lo, P.seg_hi, (int)nseg,
                                                          (int*)P.seg_order);
        k_pipe_prefill<<<(unsigned)nseg, kSerThreads, 0, s>>>(a, seg_cols, P.seg_pool, P.exact,
                                                              (unsigned long long*)P.seg_state);
        FPH_LAUNCH_CHECK();
        const size_t lds = 4 * (size_t)L.arena;
        FPH_CHECK_HIP(hipFuncSetAttribute((const void*)k_reduce_pipe,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        k_reduce_pipe<<<(unsigned)L.grid, FPH_WAVE, lds, s>>>(P);
        FPH_LAUNCH_CHECK();
        k_pipe_check<<<fph_blocks(nseg, 256), 256, 0, s>>>(P.seg_lo, P.seg_hi, P.seg_done,
                                                           (int)nseg, status);
        FPH_LAUNCH_CHECK();
    }
    k_reduce_fini<<<1, 64, 0, s>>>(status, a.pool_ptr);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
