// COO export of the filtration for external reducers (the reference's
// giotto-ph bridge, persistence.py:27-61): the upper-triangle sparse
// distance matrix with one entry per edge (i, j, length) followed by the
// zero diagonal (v, v, 0) of every vertex, written on the device in one
// launch so the host copies it back in one transfer.  Lengths are the
// correctly rounded FP32 square roots of the edges' dsq (torch.sqrt in the
// reference, persistence.py:39).
#include "fph_common.h"

namespace {

__global__ void k_coo(const int* __restrict__ ei, const int* __restrict__ ej,
                      const float* __restrict__ dsq, long long E, long long n,
                      long long* __restrict__ row, long long* __restrict__ col,
                      float* __restrict__ val) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < E) {
        row[t] = ei[t];
        col[t] = ej[t];
        val[t] = fph_sqrt_rn(dsq[t]);
    } else if (t < E + n) {
        const long long v = t - E;
        row[t] = v;
        col[t] = v;
        val[t] = 0.0f;
    }
}

}  // namespace

extern "C" int fph_coo_export(const int32_t* edge_i, const int32_t* edge_j,
                              const float* edge_dsq, int64_t E, int64_t n, int64_t* row,
                              int64_t* col, float* val, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (E < 0 || n < 0 || (E + n > 0 && (!row || !col || !val)) ||
        (E > 0 && (!edge_i || !edge_j || !edge_dsq)))
        return FPH_ERR_ARG;
    if (E + n == 0) return FPH_OK;
    k_coo<<<fph_blocks(E + n, 256), 256, 0, s>>>(edge_i, edge_j, edge_dsq, E, n,
                                                  (long long*)row, (long long*)col, val);
    FPH_LAUNCH_CHECK();
    return FPH_OK;
}
