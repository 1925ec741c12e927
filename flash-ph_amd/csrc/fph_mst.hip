// Stage (b): Boruvka minimum spanning forest and the H0 diagram.
//
// Replaces boruvka.py:17-113 and the H0 assembly of persistence.py:169-192.
// Edge weights are the filtration ranks 0..E-1 (distinct), so the forest is
// unique and equals Kruskal by rank.  The round loop is issued without host
// synchronisation: each kernel reads a device flag and returns at once after
// convergence.
#include "fph_common.h"

#include <mutex>

#include <rocprim/rocprim.hpp>

namespace {

constexpr unsigned kNone = 0xFFFFFFFFu;

struct BvkState {
    int* comp;        // current root of each vertex's component
    int* nxt;         // hook pointer of roots during a round
    unsigned* min_e;  // lightest crossing edge per root
    int* flags;       // [0] = done, [1 + r] = crossing edge seen in round r
};

__global__ void k_bvk_init(BvkState st, int n, unsigned char* __restrict__ mst, long long E,
                           int rounds) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        st.comp[i] = (int)i;
        st.nxt[i] = (int)i;
        st.min_e[i] = kNone;
    }
    if (i < E) mst[i] = 0;
    if (i < rounds + 1) st.flags[i] = 0;
}

// Every crossing edge offers itself to both endpoint components (atomicMin on
// rank); the hook pointers of the previous round are reset here.  min_e
// starts the round at kNone (written by the previous kernel) and only
// decreases, so a plain read returns a value >= the current minimum: when it
// is already <= this edge's rank, the atomic could not change anything and is
// skipped.
__global__ void k_bvk_find(BvkState st, const int* __restrict__ ei, const int* __restrict__ ej,
                           long long E, int n, int round) {
    if (st.flags[0]) return;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st.nxt[i] = (int)i;
    if (i >= E) return;
    const int a = ei[i];
    if (a < 0) return;  // unused slot of a capacity-sized edge list
    const int cu = st.comp[a], cv = st.comp[ej[i]];
    if (cu == cv) return;
    const unsigned e = (unsigned)i;
    if (st.min_e[cu] > e) atomicMin(&st.min_e[cu], e);
    if (st.min_e[cv] > e) atomicMin(&st.min_e[cv], e);
    st.flags[1 + round] = 1;
}

// Each root adds its lightest crossing edge to the forest and points at the
// component across it.  A mutual pair (both roots chose the same edge) keeps
// the smaller root, so the hook graph is a forest.
__global__ void k_bvk_hook(BvkState st, const int* __restrict__ ei, const int* __restrict__ ej,
                           int n, int round, unsigned char* __restrict__ mst) {
    if (st.flags[0]) return;
    if (!st.flags[1 + round]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st.flags[0] = 1;
        return;
    }
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || st.comp[c] != c) return;
    const unsigned e = st.min_e[c];
    if (e == kNone) return;
    mst[e] = 1;
    const int cu = st.comp[ei[e]], cv = st.comp[ej[e]];
    const int other = (cu == c) ? cv : cu;
    if (st.min_e[other] == e && other > c) return;  // mutual: c stays root
    st.nxt[c] = other;
}

// Pointer jumping with path halving: every vertex walks its root's hook chain
// to the new root.  Concurrent halving writes always store an ancestor, so
// racing readers still see a valid chain.
__global__ void k_bvk_jump(BvkState st, int n) {
    if (st.flags[0]) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int r = st.comp[v];
    // The hook graph is a forest (distinct ranks), so every walk ends at a
    // root within n steps; the bound only guards against corrupted input.
    for (int it = 0; it < n; it++) {
        const int p = st.nxt[r];
        if (p == r) break;
        const int gp = st.nxt[p];
        if (gp != p) st.nxt[r] = gp;
        r = p;
    }
    st.comp[v] = r;
    st.min_e[v] = kNone;
}

__global__ void k_bvk_minlabel_init(unsigned* __restrict__ lab, int n) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) lab[v] = kNone;
}

__global__ void k_bvk_minlabel(const int* __restrict__ comp, unsigned* __restrict__ lab, int n) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicMin(&lab[comp[v]], (unsigned)v);
}

// Component label = smallest vertex of the component (boruvka.py:91-99 hooks
// the larger root under the smaller one, so its roots are component minima).
__global__ void k_bvk_label(const int* __restrict__ comp, const unsigned* __restrict__ lab,
                            long long* __restrict__ out, int n, const int* __restrict__ flags,
                            int rounds, int* __restrict__ converged) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && out) out[v] = (long long)lab[comp[v]];
    if (v == 0 && converged) converged[0] = flags[0] ? 1 : (flags[rounds] ? 0 : 1);
}

// The whole round loop in one workgroup, its state in LDS (comp, nxt,
// min_e: 12 bytes per vertex; gfx950 gives a workgroup up to 160 KB), for
// clouds of up to kBvkLdsMaxN vertices.  Same rounds as the kernels above --
// find, hook, pointer jumping with path halving -- separated by barriers
// instead of launches, and it stops at the first round without a crossing
// edge instead of running every round: the forest (unique: distinct ranks),
// the labels and the convergence flag are those of the multi-launch form.
// Hooking goes from the edge side: an edge that is the lightest crossing
// edge of one of its two components hooks that component to the other (of a
// mutual pair only the larger root), which is the root-side rule above
// without reading the chosen edge's endpoints back from memory.  With
// CACHED, each thread keeps its (at most kBvkEdgesPerThread) edges in
// registers for all rounds, one word each (a << 16 | b: n < 2^16), and drops
// an edge once it lies inside a component, so later rounds touch LDS only
// for crossing edges.
constexpr int kBvkLdsThreads = 1024;
constexpr int kBvkLdsMaxN = 12288;          // 12 * n bytes <= 144 KB
constexpr long long kBvkLdsMaxE = 1 << 19;  // beyond, the edge sweeps favour the grid
constexpr int kBvkEdgesPerThread = 32;      // register-resident edges (CACHED)

// One edge (a, b) of rank e during find: offers e to both components.
// Returns true when the edge lies inside one component (it stays inside:
// components only merge).  min_e is reset before the find phase and only
// decreases during it, so a read never shows less than the current minimum:
// when it is already <= e the atomic is skipped (offers to one large
// component would otherwise serialise on its LDS word).
__device__ __forceinline__ bool bvk_offer(const int* comp, unsigned* min_e, int a, int b,
                                          unsigned e, int& cross) {
    if (a < 0) return true;  // unused slot of a capacity-sized edge list
    const int cu = comp[a], cv = comp[b];
    if (cu == cv) return true;
    if (min_e[cu] > e) atomicMin(&min_e[cu], e);
    if (min_e[cv] > e) atomicMin(&min_e[cv], e);
    cross = 1;
    return false;
}

// One edge during hook: hooks a component that chose it (the forest edges
// themselves are marked from the roots' side, mst[min_e[root]]).
__device__ __forceinline__ void bvk_hook_edge(const int* comp, int* nxt, const unsigned* min_e,
                                              int a, int b, unsigned e) {
    if (a < 0) return;
    const int cu = comp[a], cv = comp[b];
    if (cu == cv) return;
    const bool mu = min_e[cu] == e, mv = min_e[cv] == e;
    if (mu && (!mv || cu > cv)) nxt[cu] = cv;
    if (mv && (!mu || cv > cu)) nxt[cv] = cu;
}

template <bool CACHED>
__global__ __launch_bounds__(kBvkLdsThreads) void k_bvk_lds(
    const int* __restrict__ ei, const int* __restrict__ ej, long long E, int n, int rounds,
    unsigned char* __restrict__ mst, long long* __restrict__ comp_label,
    int* __restrict__ converged) {
    extern __shared__ int lds[];
    int* comp = lds;
    int* nxt = lds + n;
    unsigned* min_e = reinterpret_cast<unsigned*>(lds + 2 * n);
    __shared__ int s_cross;
    const int tid = threadIdx.x;
    unsigned ce[CACHED ? kBvkEdgesPerThread : 1];
    if constexpr (CACHED) {
#pragma unroll
        for (int u = 0; u < kBvkEdgesPerThread; u++) {
            const long long e = tid + (long long)u * kBvkLdsThreads;
            const int a = e < E ? ei[e] : -1;
            ce[u] = a < 0 ? 0xFFFFFFFFu : ((unsigned)a << 16 | (unsigned)ej[e]);
        }
    }
    for (int v = tid; v < n; v += kBvkLdsThreads) {
        comp[v] = v;
        nxt[v] = v;
        min_e[v] = kNone;
    }
    for (long long e = tid; e < E; e += kBvkLdsThreads) mst[e] = 0;
    int conv = 0;
    for (int r = 0; r < rounds; r++) {
        if (tid == 0) s_cross = 0;
        __syncthreads();
        // find: every crossing edge offers itself to both components
        int cross = 0;
        if constexpr (CACHED) {
#pragma unroll
            for (int u = 0; u < kBvkEdgesPerThread; u++)
                if (ce[u] != 0xFFFFFFFFu &&
                    bvk_offer(comp, min_e, (int)(ce[u] >> 16), (int)(ce[u] & 0xFFFFu),
                              (unsigned)(tid + u * kBvkLdsThreads), cross))
                    ce[u] = 0xFFFFFFFFu;
        } else {
            for (long long e = tid; e < E; e += kBvkLdsThreads)
                (void)bvk_offer(comp, min_e, ei[e], ej[e], (unsigned)e, cross);
        }
        if (cross) s_cross = 1;
        __syncthreads();
        if (!s_cross) {
            conv = 1;
            break;
        }
        // hook, from the edge side
        if constexpr (CACHED) {
#pragma unroll
            for (int u = 0; u < kBvkEdgesPerThread; u++)
                if (ce[u] != 0xFFFFFFFFu)
                    bvk_hook_edge(comp, nxt, min_e, (int)(ce[u] >> 16), (int)(ce[u] & 0xFFFFu),
                                  (unsigned)(tid + u * kBvkLdsThreads));
        } else {
            for (long long e = tid; e < E; e += kBvkLdsThreads)
                bvk_hook_edge(comp, nxt, min_e, ei[e], ej[e], (unsigned)e);
        }
        for (int c = tid; c < n; c += kBvkLdsThreads)
            if (comp[c] == c && min_e[c] != kNone) mst[min_e[c]] = 1;
        __syncthreads();
        // jump: every vertex walks its root's hook chain to the new root
        for (int v = tid; v < n; v += kBvkLdsThreads) {
            int x = comp[v];
            for (int it = 0; it < n; it++) {
                const int q = nxt[x];
                if (q == x) break;
                const int gq = nxt[q];
                if (gq != q) nxt[x] = gq;
                x = q;
            }
            comp[v] = x;
        }
        __syncthreads();
        for (int v = tid; v < n; v += kBvkLdsThreads) {
            nxt[v] = v;
            min_e[v] = kNone;
        }
    }
    if (comp_label) {
        // label = smallest vertex of the component (nxt reused as labels)
        unsigned* lab = reinterpret_cast<unsigned*>(nxt);
        __syncthreads();
        for (int v = tid; v < n; v += kBvkLdsThreads) lab[v] = kNone;
        __syncthreads();
        for (int v = tid; v < n; v += kBvkLdsThreads) atomicMin(&lab[comp[v]], (unsigned)v);
        __syncthreads();
        for (int v = tid; v < n; v += kBvkLdsThreads) comp_label[v] = (long long)lab[comp[v]];
    }
    if (tid == 0 && converged) converged[0] = conv;
}

__global__ void k_u8_to_i32(const unsigned char* __restrict__ m, long long E, int* __restrict__ o) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) o[e] = m[e];
}

// H0 rows: (0, sqrt(dsq)) for forest edges in rank order, then (0, inf) for
// each component (persistence.py:172-188).  Exactly n rows.
__global__ void k_h0_write(const unsigned char* __restrict__ mst, const int* __restrict__ pos,
                           const float* __restrict__ dsq, long long E, long long n,
                           float* __restrict__ h0, long long* __restrict__ mst_idx,
                           long long* __restrict__ n_mst) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nm = E > 0 ? (long long)pos[E - 1] + mst[E - 1] : 0;
    if (i < E && mst[i] && pos[i] < n) {
        const long long p = pos[i];
        h0[2 * p] = 0.0f;
        h0[2 * p + 1] = fph_sqrt_rn(dsq[i]);
        if (mst_idx) mst_idx[p] = i;
    }
    if (i < n && i >= nm) {
        h0[2 * i] = 0.0f;
        h0[2 * i + 1] = __builtin_inff();
    }
    if (i == 0 && n_mst) n_mst[0] = nm;
}

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Exclusive scan of the forest mask in three launches and no memset (for
// graph capture): per-block counts, one-block scan of the counts, per-block
// scan plus offset.
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int block_excl_sum(int v, int* sh, int* total) {
    // 1024 threads = 16 wavefronts: wave scans, then a scan of the wave sums
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (threadIdx.x < 16) {
        int t = sh[threadIdx.x];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(t, o, 16);
            if ((int)threadIdx.x >= o) t += y;
        }
        sh[16 + threadIdx.x] = t;
    }
    __syncthreads();
    const int before = wv ? sh[16 + wv - 1] : 0;
    if (total) *total = sh[31];
    const int r = before + x - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanBlock) void k_mask_block_count(
    const unsigned char* __restrict__ m, long long E, int* __restrict__ bsum) {
    __shared__ int sh[32];
    const long long i = (long long)blockIdx.x * kScanBlock + threadIdx.x;
    const int v = i < E ? (int)m[i] : 0;
    int tot;
    block_excl_sum(v, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_block_counts(int* __restrict__ bsum, int nb) {
    __shared__ int sh[32];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanBlock) {
        const int b = b0 + (int)threadIdx.x;
        const int v = b < nb ? bsum[b] : 0;
        int tot;
        const int ex = block_excl_sum(v, sh, &tot);
        if (b < nb) bsum[b] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_mask_block_scan(
    const unsigned char* __restrict__ m, long long E, const int* __restrict__ bsum,
    int* __restrict__ pos) {
    __shared__ int sh[32];
    const long long i = (long long)blockIdx.x * kScanBlock + threadIdx.x;
    const int v = i < E ? (int)m[i] : 0;
    const int ex = block_excl_sum(v, sh, nullptr);
    if (i < E) pos[i] = bsum[blockIdx.x] + ex;
}

}  // namespace

// Capacity form of fph_h0_diagram for graph capture: E = cap slots (unused
// slots have mask 0), no memset nodes.
int fph_h0_diagram_capacity(const uint8_t* mst_mask, const float* edge_dsq, int64_t cap,
                            int64_t n, float* h0, void* ws, size_t* ws_bytes, hipStream_t s) {
    if (cap < 1 || n < 1 || !ws_bytes) return FPH_ERR_ARG;
    const long long nb = (cap + kScanBlock - 1) / kScanBlock;
    const size_t need = align_up(4 * (size_t)nb) + align_up(4 * (size_t)cap);
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    int* bsum = (int*)ws;
    int* pos = (int*)((char*)ws + align_up(4 * (size_t)nb));
    k_mask_block_count<<<(unsigned)nb, kScanBlock, 0, s>>>(mst_mask, cap, bsum);
    k_scan_block_counts<<<1, kScanBlock, 0, s>>>(bsum, (int)nb);
    k_mask_block_scan<<<(unsigned)nb, kScanBlock, 0, s>>>(mst_mask, cap, bsum, pos);
    FPH_LAUNCH_CHECK();
    const long long span = std::max<long long>(cap, n);
    k_h0_write<<<fph_blocks(span, 256), 256, 0, s>>>(mst_mask, pos, edge_dsq, cap, n, h0, nullptr,
                                                    nullptr);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

// Allows the one-workgroup form its LDS (once per process; called before
// any stream capture by the graph builder, since it is not a stream call).
hipError_t fph_mst_prepare() {
    static std::once_flag once;
    static hipError_t rc = hipSuccess;
    std::call_once(once, [] {
        rc = hipFuncSetAttribute((const void*)k_bvk_lds<true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 12 * kBvkLdsMaxN);
        if (rc == hipSuccess)
            rc = hipFuncSetAttribute((const void*)k_bvk_lds<false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     12 * kBvkLdsMaxN);
    });
    return rc;
}

extern "C" int fph_boruvka_rounds(int64_t n) {
    int64_t m = n < 2 ? 2 : n;
    int r = 0;
    while ((1LL << r) < m) r++;
    return r + 2;
}

extern "C" size_t fph_boruvka_scratch_bytes(int64_t n) {
    const int R = fph_boruvka_rounds(n);
    return 4 * align_up(4 * (size_t)std::max<int64_t>(n, 1)) + align_up(4 * (size_t)(R + 2));
}

// Issues the whole Boruvka forest on a stream: the one-workgroup form for
// small complexes (no scratch), else the round kernels with their state in
// `w` (fph_boruvka_scratch_bytes(n) bytes, device memory).
static int bvk_issue(const int32_t* edge_i, const int32_t* edge_j, int64_t E, int64_t n,
                     uint8_t* mst_mask, int64_t* comp_label, int32_t* converged, char* w,
                     hipStream_t s) {
    const int R = fph_boruvka_rounds(n);
    if (n <= kBvkLdsMaxN && E <= kBvkLdsMaxE) {
        const int lds_b = 12 * (int)n;
        FPH_CHECK_HIP(fph_mst_prepare());
        if (E <= (long long)kBvkEdgesPerThread * kBvkLdsThreads)
            k_bvk_lds<true><<<1, kBvkLdsThreads, lds_b, s>>>(edge_i, edge_j, E, (int)n, R,
                                                            mst_mask, (long long*)comp_label,
                                                            converged);
        else
            k_bvk_lds<false><<<1, kBvkLdsThreads, lds_b, s>>>(edge_i, edge_j, E, (int)n, R,
                                                             mst_mask, (long long*)comp_label,
                                                             converged);
        FPH_LAUNCH_CHECK();
        return FPH_OK;
    }
    const size_t vb = align_up(4 * (size_t)n);
    BvkState st;
    st.comp = (int*)w;
    st.nxt = (int*)(w + vb);
    st.min_e = (unsigned*)(w + 2 * vb);
    unsigned* lab = (unsigned*)(w + 3 * vb);
    st.flags = (int*)(w + 4 * vb);
    // k_bvk_init sets comp, nxt, min_e and every round flag; lab is set by
    // k_bvk_minlabel_init: nothing in `w` needs to start at zero.
    const long long span = std::max<long long>(std::max<long long>(E, n), R + 1);
    k_bvk_init<<<fph_blocks(span, 256), 256, 0, s>>>(st, (int)n, mst_mask, E, R);
    FPH_LAUNCH_CHECK();
    const long long fspan = std::max<long long>(E, n);
    for (int r = 0; r < R; r++) {
        k_bvk_find<<<fph_blocks(fspan, 256), 256, 0, s>>>(st, edge_i, edge_j, E, (int)n, r);
        k_bvk_hook<<<fph_blocks(n, 256), 256, 0, s>>>(st, edge_i, edge_j, (int)n, r, mst_mask);
        k_bvk_jump<<<fph_blocks(n, 256), 256, 0, s>>>(st, (int)n);
    }
    FPH_LAUNCH_CHECK();
    if (comp_label) {
        k_bvk_minlabel_init<<<fph_blocks(n, 256), 256, 0, s>>>(lab, (int)n);
        k_bvk_minlabel<<<fph_blocks(n, 256), 256, 0, s>>>(st.comp, lab, (int)n);
    }
    k_bvk_label<<<fph_blocks(n, 256), 256, 0, s>>>(st.comp, lab, (long long*)comp_label, (int)n,
                                                  st.flags, R, converged);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}

extern "C" int fph_boruvka_mst(const int32_t* edge_i, const int32_t* edge_j, int64_t E,
                               int64_t n, uint8_t* mst_mask, int64_t* comp_label,
                               int32_t* converged, void* scratch, size_t scratch_bytes,
                               void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || n < 1 || !scratch) return FPH_ERR_ARG;
    if (n > 0x7FFFFFFFLL || E >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    if (scratch_bytes < fph_boruvka_scratch_bytes(n)) return FPH_ERR_ARG;
    return bvk_issue(edge_i, edge_j, E, n, mst_mask, comp_label, converged, (char*)scratch, s);
}

// ---- allocation-free Boruvka ------------------------------------------------
//
// The round kernels' state for up to FPH_BORUVKA_NOALLOC_MAX_N vertices lives
// in this file-scope array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that address is used.
namespace {
constexpr size_t kNoallocBytes =
    4 * (4 * (size_t)FPH_BORUVKA_NOALLOC_MAX_N) + 256;  // comp, nxt, min_e, lab + flags
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kMaxDevices];

// Device address of `scratch` on `dev` (looked up once per device).
static int noalloc_scratch(int dev, char** out) {
    if (dev < 0 || dev >= kMaxDevices) return FPH_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        int cur = -1;
        FPH_CHECK_HIP(hipGetDevice(&cur));
        if (cur != dev) FPH_CHECK_HIP(hipSetDevice(dev));
        void* p = nullptr;
        hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) {
            hipError_t rc2 = hipSetDevice(cur);
            if (rc == hipSuccess) rc = rc2;
        }
        if (rc != hipSuccess || !p) return FPH_ERR_HIP;
        g_noalloc_ptr[dev] = p;
    }
    *out = (char*)g_noalloc_ptr[dev];
    return FPH_OK;
}

extern "C" int fph_boruvka_mst_noalloc(const int32_t* edge_i, const int32_t* edge_j, int64_t E,
                                       int64_t n, uint8_t* mst_mask, int64_t* comp_label,
                                       int32_t* converged, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || n < 1) return FPH_ERR_ARG;
    if (n > FPH_BORUVKA_NOALLOC_MAX_N || E >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    if (fph_boruvka_scratch_bytes(n) > kNoallocBytes) return FPH_ERR_LIMIT;
    int dev = -1;
    if (s) FPH_CHECK_HIP(hipStreamGetDevice(s, &dev));
    else FPH_CHECK_HIP(hipGetDevice(&dev));
    char* w = nullptr;
    const int rc = noalloc_scratch(dev, &w);
    if (rc != FPH_OK) return rc;
    return bvk_issue(edge_i, edge_j, E, n, mst_mask, comp_label, converged, w, s);
}

extern "C" int fph_h0_diagram(const uint8_t* mst_mask, const float* edge_dsq, int64_t E,
                              int64_t n, float* h0, int64_t* mst_idx, int64_t* n_mst,
                              void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || n < 1 || !ws_bytes) return FPH_ERR_ARG;
    size_t tmp = 0;
    FPH_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp, (const int*)nullptr, (int*)nullptr, 0,
                                          (size_t)std::max<int64_t>(E, 1), rocprim::plus<int>(), s));
    tmp = align_up(tmp);
    const size_t need = tmp + 2 * align_up(4 * (size_t)std::max<int64_t>(E, 1));
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w = (char*)ws;
    int* m32 = (int*)(w + tmp);
    int* pos = (int*)(w + tmp + align_up(4 * (size_t)std::max<int64_t>(E, 1)));
    if (E > 0) {
        k_u8_to_i32<<<fph_blocks(E, 256), 256, 0, s>>>(mst_mask, E, m32);
        FPH_LAUNCH_CHECK();
        size_t t1 = tmp;
        FPH_CHECK_HIP(rocprim::exclusive_scan(ws, t1, m32, pos, 0, (size_t)E, rocprim::plus<int>(), s));
    }
    const long long span = std::max<long long>(E, n);
    k_h0_write<<<fph_blocks(span, 256), 256, 0, s>>>(mst_mask, pos, edge_dsq, E, n, h0,
                                                    (long long*)mst_idx, (long long*)n_mst);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
