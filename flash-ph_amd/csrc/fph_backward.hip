// Gradient of persistence diagrams with respect to the points.
//
// Every finite birth or death value of a Vietoris-Rips diagram is the length
// |x_u - x_v| of one edge (the diameter edge of the critical simplex), so
// d value / d x_u = (x_u - x_v) / |x_u - x_v| and the opposite for x_v.
// Given the upstream gradient of every (row, birth/death) entry and the edge
// behind it, the gradient of the points is
//   grad x_v = sum over edges e = (v, u):  w_e (x_v - x_u),
//   w_e = (sum of the upstream gradients of the entries on e) / |e|.
// Both sums run in a fixed order (entries sorted stably by edge; a vertex's
// edges in CSR order), so the result does not depend on scheduling.
// The reference has no gradient path; this is new in this library.
#include "fph_common.h"

#include <rocprim/rocprim.hpp>

namespace {

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

__global__ void k_contrib_keys(const int* __restrict__ edge, long long M,
                               unsigned* __restrict__ keys, unsigned* __restrict__ idx) {
    long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int e = edge[m];
    keys[m] = e < 0 ? 0xFFFFFFFFu : (unsigned)e;  // entries without an edge sort last
    idx[m] = (unsigned)m;
}

__device__ __forceinline__ long long lb_keys(const unsigned* k, long long n, unsigned v) {
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (k[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// w_e = (sum of the entries on e, in entry order) / |e|; 0 when |e| = 0
// (the subgradient 0 of |x| at 0) or when no entry uses e.
__global__ void k_edge_weight(const unsigned* __restrict__ skeys,
                              const unsigned* __restrict__ sidx, long long M,
                              const float* __restrict__ val, const float* __restrict__ filt,
                              long long E, float* __restrict__ w) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const long long lo = lb_keys(skeys, M, (unsigned)e);
    const long long hi = lb_keys(skeys, M, (unsigned)e + 1u);
    double g = 0.0;
    for (long long m = lo; m < hi; m++) g += (double)val[sidx[m]];
    const float f = filt[e];
    w[e] = (hi > lo && f > 0.0f) ? (float)(g / (double)f) : 0.0f;
}

__device__ __forceinline__ int cloud_of(const long long* vbase, int C, long long v) {
    int lo = 0, hi = C;  // last c with vbase[c] <= v
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (vbase[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One thread per vertex: grad x_v = sum over its CSR edges of w_e (x_v - x_u).
__global__ void k_vertex_grad(const float* __restrict__ pts, const long long* __restrict__ foff,
                              const int* __restrict__ dims, const long long* __restrict__ vbase,
                              int C, long long N, const int* __restrict__ adj_ptr,
                              const int* __restrict__ adj_idx, const int* __restrict__ adj_rank,
                              const float* __restrict__ w, float* __restrict__ grad) {
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int c = cloud_of(vbase, C, v);
    const int d = dims[c];
    const float* P = pts + foff[c];
    const long long vb = vbase[c];
    const float* xv = P + (v - vb) * d;
    float* gv = grad + foff[c] + (v - vb) * d;
    for (int k = 0; k < d; k++) gv[k] = 0.0f;
    for (int a = adj_ptr[v]; a < adj_ptr[v + 1]; a++) {
        const float we = w[adj_rank[a]];
        if (we == 0.0f) continue;
        const float* xu = P + ((long long)adj_idx[a] - vb) * d;
        for (int k = 0; k < d; k++) gv[k] += we * (xv[k] - xu[k]);
    }
}

}  // namespace

extern "C" int fph_backward_scatter(const float* pts, const int64_t* foff, const int32_t* dims,
                                    const int64_t* vbase, int64_t C, int64_t N,
                                    const float* edge_filt, int64_t E, const int32_t* adj_ptr,
                                    const int32_t* adj_idx, const int32_t* adj_rank,
                                    const int32_t* contrib_edge, const float* contrib_val,
                                    int64_t M, float* grad_pts, void* ws, size_t* ws_bytes,
                                    void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (C < 1 || N < 0 || E < 0 || M < 0 || !ws_bytes) return FPH_ERR_ARG;
    if (M >= 0x7FFFFFFFLL || E >= 0x7FFFFFFFLL) return FPH_ERR_LIMIT;
    const size_t m4 = align_up(4 * (size_t)std::max<int64_t>(M, 1));
    const size_t e4 = align_up(4 * (size_t)std::max<int64_t>(E, 1));
    size_t tmp = 0;
    FPH_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmp, (const unsigned*)nullptr,
                                            (unsigned*)nullptr, (const unsigned*)nullptr,
                                            (unsigned*)nullptr,
                                            (unsigned)std::max<int64_t>(M, 1), 0, 32, s));
    tmp = align_up(tmp);
    const size_t need = tmp + 4 * m4 + e4;
    if (!ws) { *ws_bytes = need; return FPH_OK; }
    if (*ws_bytes < need) return FPH_ERR_ARG;
    char* w8 = (char*)ws;
    unsigned* keys = (unsigned*)(w8 + tmp);
    unsigned* idx = (unsigned*)(w8 + tmp + m4);
    unsigned* skeys = (unsigned*)(w8 + tmp + 2 * m4);
    unsigned* sidx = (unsigned*)(w8 + tmp + 3 * m4);
    float* w = (float*)(w8 + tmp + 4 * m4);
    if (M > 0) {
        k_contrib_keys<<<fph_blocks(M, 256), 256, 0, s>>>(contrib_edge, M, keys, idx);
        FPH_LAUNCH_CHECK();
        size_t t = tmp;
        FPH_CHECK_HIP(rocprim::radix_sort_pairs(ws, t, keys, skeys, idx, sidx, (unsigned)M, 0,
                                                32, s));
    }
    if (E > 0) {
        k_edge_weight<<<fph_blocks(E, 256), 256, 0, s>>>(skeys, sidx, M, contrib_val, edge_filt,
                                                        E, w);
        FPH_LAUNCH_CHECK();
    }
    if (N > 0) {
        k_vertex_grad<<<fph_blocks(N, 128), 128, 0, s>>>(
            pts, (const long long*)foff, dims, (const long long*)vbase, (int)C, N, adj_ptr,
            adj_idx, adj_rank, w, grad_pts);
        FPH_LAUNCH_CHECK();
    }
    return FPH_OK;
}
