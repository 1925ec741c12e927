// H0 straight from the screened keys, without sorting the edges: the
// persistence of dimension 0 needs only the minimum spanning forest, and the
// packed key of an edge (float_bits(dsq) << 32 | j (j - 1) / 2 + i, rips.py:
// 41-49) is unique and decodes back to (i, j) while n <= 92682 (the colex
// field does not wrap).  The forest under key order is the forest under the
// reference's rank order (rank = position in key order, boruvka.py:17-113),
// so Boruvka runs on the keys themselves with 64-bit minima per component,
// and only its n - 1 (at most) edges are sorted for the diagram
// (persistence.py:160-186).  The edge list is never sorted or turned into a
// CSR, which is what lets H0 take clouds whose edge count passes 2^31 (the
// reference's limit, kernels/edge_kernel.py:203-205): 2^31 edges are 16 GiB
// of keys, resident in HBM, read once per round.
#include "fph_common.h"

#include <rocprim/rocprim.hpp>

namespace {

constexpr unsigned long long kNoKey = ~0ull;
constexpr int kThreads = 256;
constexpr unsigned kEdgeBlocks = 4096;  // grid-stride over the edges

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

struct HkState {
    int* comp;                  // current root of each vertex
    int* nxt;                   // hook pointer of roots during a round
    unsigned long long* best;   // lightest crossing key per root
    int* flags;                 // [0] done, [1 + r] crossing key seen in round r
    unsigned long long* forest; // forest keys (append)
    unsigned long long* nf;     // forest size
};

// (i, j) of colex index c = j (j - 1) / 2 + i, 0 <= i < j.
__device__ __forceinline__ void colex_pair(unsigned long long c, int* i, int* j) {
    long long jj = (long long)((1.0 + sqrt(1.0 + 8.0 * (double)c)) * 0.5);
    while (jj * (jj - 1) / 2 > (long long)c) jj--;
    while ((jj + 1) * jj / 2 <= (long long)c) jj++;
    *j = (int)jj;
    *i = (int)((long long)c - jj * (jj - 1) / 2);
}

__global__ void k_hk_init(HkState st, int n, int rounds) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) {
        st.comp[v] = v;
        st.nxt[v] = v;
        st.best[v] = kNoKey;
        st.forest[v] = kNoKey;  // unused slots sort last
    }
    if (v < rounds + 1) st.flags[v] = 0;
    if (v == 0) st.nf[0] = 0ull;
}

// Every crossing edge offers its key to both endpoint components.  best
// starts the round at kNoKey and only decreases, so a plain read >= the
// current minimum: the atomic is skipped when it could not lower it.
__global__ void k_hk_find(HkState st, const unsigned long long* __restrict__ keys,
                          unsigned long long E, int round) {
    if (st.flags[0]) return;
    bool crossing = false;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += stride) {
        const unsigned long long key = keys[e];
        int i, j;
        colex_pair(key & 0xFFFFFFFFull, &i, &j);
        const int cu = st.comp[i], cv = st.comp[j];
        if (cu == cv) continue;
        crossing = true;
        if (st.best[cu] > key) atomicMin(&st.best[cu], key);
        if (st.best[cv] > key) atomicMin(&st.best[cv], key);
    }
    if (__ballot(crossing) && fph_lane() == 0) st.flags[1 + round] = 1;
}

// Each root takes its lightest crossing edge into the forest and hooks to
// the component across it; of a mutual pair (both chose one edge) the
// smaller root stays a root and records the edge, so each forest edge is
// recorded once and the hook graph is a forest.
__global__ void k_hk_hook(HkState st, int n, int round) {
    if (st.flags[0]) return;
    if (!st.flags[1 + round]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st.flags[0] = 1;
        return;
    }
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || st.comp[c] != c) return;
    const unsigned long long key = st.best[c];
    if (key == kNoKey) return;
    int i, j;
    colex_pair(key & 0xFFFFFFFFull, &i, &j);
    const int cu = st.comp[i], cv = st.comp[j];
    const int other = (cu == c) ? cv : cu;
    const bool mutual = st.best[other] == key;
    if (mutual && other > c) {
        st.forest[atomicAdd(st.nf, 1ull)] = key;  // c stays root
        return;
    }
    if (!mutual) st.forest[atomicAdd(st.nf, 1ull)] = key;
    st.nxt[c] = other;
}

// Pointer jumping with path halving (the hook graph is a forest).
__global__ void k_hk_jump(HkState st, int n) {
    if (st.flags[0]) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int r = st.comp[v];
    for (int it = 0; it < n; it++) {
        const int p = st.nxt[r];
        if (p == r) break;
        const int gp = st.nxt[p];
        if (gp != p) st.nxt[r] = gp;
        r = p;
    }
    st.comp[v] = r;
    st.best[v] = kNoKey;
}

// H0 rows: the sorted forest keys' lengths, then (0, inf) per component.
__global__ void k_hk_rows(const unsigned long long* __restrict__ sorted,
                          const unsigned long long* __restrict__ nf, int n,
                          const int* __restrict__ flags, int rounds, float* __restrict__ h0,
                          int* __restrict__ converged) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long m = nf[0];
    if (k < n) {
        float death = __builtin_inff();
        if ((unsigned long long)k < m) {
            const unsigned bits = (unsigned)(sorted[k] >> 32);
            float dsq;
            memcpy(&dsq, &bits, 4);
            death = fph_sqrt_rn(dsq);
        }
        h0[2 * k] = 0.0f;
        h0[2 * k + 1] = death;
    }
    if (k == 0 && converged) converged[0] = flags[0] ? 1 : (flags[rounds] ? 0 : 1);
}

int rounds_for(int64_t n) {
    int64_t m = n < 2 ? 2 : n;
    int r = 0;
    while ((1LL << r) < m) r++;
    return r + 2;
}

}  // namespace

extern "C" int fph_h0_from_keys(const uint64_t* keys, int64_t E, int64_t n, float* h0,
                                int32_t* converged, void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || n < 1 || !ws_bytes || (E > 0 && !keys)) return FPH_ERR_ARG;
    if (n > 92682) return FPH_ERR_LIMIT;  // the colex field must not wrap
    const int R = rounds_for(n);
    const size_t v4 = align_up(4 * (size_t)n), v8 = align_up(8 * (size_t)n);
    size_t tmp = 0;
    FPH_CHECK_HIP(rocprim::radix_sort_keys(nullptr, tmp, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (size_t)n, 0, 64, s));
    tmp = align_up(tmp);
    const size_t need = 2 * v4 + 3 * v8 + align_up(4 * (size_t)(R + 2)) + 256 + tmp;
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need || !h0) return FPH_ERR_ARG;
    char* w = (char*)ws;
    HkState st;
    st.comp = (int*)w;
    st.nxt = (int*)(w + v4);
    st.best = (unsigned long long*)(w + 2 * v4);
    st.forest = (unsigned long long*)(w + 2 * v4 + v8);
    unsigned long long* sorted = (unsigned long long*)(w + 2 * v4 + 2 * v8);
    st.flags = (int*)(w + 2 * v4 + 3 * v8);
    st.nf = (unsigned long long*)(w + 2 * v4 + 3 * v8 + align_up(4 * (size_t)(R + 2)));
    void* sort_tmp = w + 2 * v4 + 3 * v8 + align_up(4 * (size_t)(R + 2)) + 256;
    const unsigned vb = fph_blocks(std::max<int64_t>(n, R + 1), kThreads);
    k_hk_init<<<vb, kThreads, 0, s>>>(st, (int)n, R);
    FPH_LAUNCH_CHECK();
    const unsigned eb = (unsigned)std::max<long long>(
        1, std::min<long long>(kEdgeBlocks, (E + kThreads - 1) / kThreads));
    for (int r = 0; r < R; r++) {
        if (E > 0)
            k_hk_find<<<eb, kThreads, 0, s>>>(st, (const unsigned long long*)keys,
                                               (unsigned long long)E, r);
        k_hk_hook<<<fph_blocks(n, kThreads), kThreads, 0, s>>>(st, (int)n, r);
        k_hk_jump<<<fph_blocks(n, kThreads), kThreads, 0, s>>>(st, (int)n);
    }
    FPH_LAUNCH_CHECK();
    // at most n - 1 forest keys, the unused slots kNoKey: one sort of n keys
    size_t t = tmp;
    FPH_CHECK_HIP(rocprim::radix_sort_keys(sort_tmp, t, st.forest, sorted, (size_t)n, 0, 64, s));
    k_hk_rows<<<fph_blocks(n, kThreads), kThreads, 0, s>>>(sorted, st.nf, (int)n, st.flags, R, h0,
                                                          converged);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
