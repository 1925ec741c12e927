// Graph-captured H0 pipeline: screen -> sort -> vertex CSR -> Boruvka ->
// H0 diagram as one hipGraph, replayed for every new cloud of the same shape.
//
// The eager pipeline reads the edge count back to size the next stage.
// Here every stage is sized by a capacity `cap` fixed at capture time; the
// live count stays on the device (count_dev[0]), unused slots carry
// all-ones sentinel keys that sort last and that the CSR and Boruvka
// kernels skip, and the sorts and scans record kernels only (no memset
// node, no host synchronisation).  The caller checks count_dev[0] <= cap
// after a replay and re-captures with a larger capacity otherwise; results
// are then identical to the eager entry points'.
// Replaces, for repeated calls, the host-driven stage sequence of
// persistence.py:140-186 (rips_filtration, gpu_boruvka_mst, H0 assembly).
#include "fph_common.h"

namespace {

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

struct GraphSizes {
    size_t kv, mfma, edges, bvk, h0, total;
};

int pipeline_sizes(const float* pts, int64_t n, int d, float thr_sq, int screen, int64_t cap,
                   GraphSizes* z, hipStream_t s) {
    z->kv = align_up(8 * (size_t)cap);
    z->mfma = 0;
    if (screen == 1) {
        uint64_t cnt[2];  // only a size query: nothing is launched
        int rc = fph_edge_screen_mfma(pts, n, d, thr_sq, nullptr, nullptr, cap, cnt, cnt + 1,
                                      nullptr, &z->mfma, s);
        if (rc) return rc;
        z->mfma = align_up(z->mfma);
    }
    int rc = fph_edges_csr_capacity(nullptr, nullptr, nullptr, cap, n, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                    &z->edges, s);
    if (rc) return rc;
    z->edges = align_up(z->edges);
    z->bvk = align_up(fph_boruvka_scratch_bytes(n));
    rc = fph_h0_diagram_capacity(nullptr, nullptr, cap, n, nullptr, nullptr, &z->h0, s);
    if (rc) return rc;
    z->h0 = align_up(z->h0);
    z->total = 2 * z->kv + z->mfma + z->edges + z->bvk + z->h0;
    return FPH_OK;
}

struct Graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

}  // namespace

extern "C" int fph_h0_pipeline(const float* pts, int64_t n, int d, float thr_sq, int screen,
                               int64_t cap, uint64_t* count_dev, int32_t* edge_i,
                               int32_t* edge_j, float* edge_dsq, float* edge_filt,
                               int32_t* adj_ptr, int32_t* adj_idx, int32_t* adj_rank,
                               float* adj_dsq, uint8_t* mst_mask, float* h0,
                               int32_t* converged, void* ws, size_t* ws_bytes, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (n < 2 || d < 1 || cap < 1 || (screen != 0 && screen != 1) || !ws_bytes)
        return FPH_ERR_ARG;
    if (n > 92682) return FPH_ERR_LIMIT;
    GraphSizes z;
    int rc = pipeline_sizes(pts, n, d, thr_sq, screen, cap, &z, s);
    if (rc) return rc;
    if (!ws) { *ws_bytes = z.total; return FPH_OK; }
    if (*ws_bytes < z.total || !count_dev || !converged) return FPH_ERR_ARG;
    char* w = (char*)ws;
    uint64_t* keys = (uint64_t*)w;
    uint64_t* vals = (uint64_t*)(w + z.kv);
    char* wm = w + 2 * z.kv;
    char* we = wm + z.mfma;
    char* wb = we + z.edges;
    char* wh = wb + z.bvk;
    if (screen == 1) {
        size_t b = z.mfma;
        rc = fph_edge_screen_mfma(pts, n, d, thr_sq, keys, vals, cap, count_dev, count_dev + 1,
                                  wm, &b, stream);
    } else {
        rc = fph_edge_screen(pts, n, d, thr_sq, keys, vals, cap, count_dev, stream);
    }
    if (rc) return rc;
    size_t be = z.edges;
    rc = fph_edges_csr_capacity(keys, vals, count_dev, cap, n, edge_i, edge_j, edge_dsq,
                                edge_filt, adj_ptr, adj_idx, adj_rank, adj_dsq, we, &be, s);
    if (rc) return rc;
    rc = fph_boruvka_mst(edge_i, edge_j, cap, n, mst_mask, nullptr, converged, wb, z.bvk, stream);
    if (rc) return rc;
    size_t bh = z.h0;
    return fph_h0_diagram_capacity(mst_mask, edge_dsq, cap, n, h0, wh, &bh, s);
}

extern "C" int fph_h0_graph_create(const float* pts, int64_t n, int d, float thr_sq, int screen,
                                   int64_t cap, uint64_t* count_dev, int32_t* edge_i,
                                   int32_t* edge_j, float* edge_dsq, float* edge_filt,
                                   int32_t* adj_ptr, int32_t* adj_idx, int32_t* adj_rank,
                                   float* adj_dsq, uint8_t* mst_mask, float* h0,
                                   int32_t* converged, void* ws, size_t ws_bytes,
                                   void** graph_out) {
    if (!ws || !graph_out) return FPH_ERR_ARG;
    *graph_out = nullptr;
    FPH_CHECK_HIP(fph_mst_prepare());  // not a stream call: before the capture
    // capture on a private stream (the legacy default stream cannot be
    // captured); the graph is launched later on the caller's stream
    hipStream_t cs;
    FPH_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return FPH_ERR_HIP;
    }
    size_t b = ws_bytes;
    const int rc = fph_h0_pipeline(pts, n, d, thr_sq, screen, cap, count_dev, edge_i, edge_j,
                                   edge_dsq, edge_filt, adj_ptr, adj_idx, adj_rank, adj_dsq,
                                   mst_mask, h0, converged, ws, &b, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != FPH_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return rc != FPH_OK ? rc : FPH_ERR_HIP;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return FPH_ERR_HIP;
    }
    *graph_out = new Graph{g, x};
    return FPH_OK;
}

extern "C" int fph_graph_launch(void* graph, void* stream) {
    if (!graph) return FPH_ERR_ARG;
    FPH_CHECK_HIP(hipGraphLaunch(static_cast<Graph*>(graph)->exec, (hipStream_t)stream));
    return FPH_OK;
}

extern "C" int fph_graph_destroy(void* graph) {
    if (!graph) return FPH_OK;
    Graph* g = static_cast<Graph*>(graph);
    const hipError_t a = hipGraphExecDestroy(g->exec);
    const hipError_t b = hipGraphDestroy(g->graph);
    delete g;
    return (a == hipSuccess && b == hipSuccess) ? FPH_OK : FPH_ERR_HIP;
}
