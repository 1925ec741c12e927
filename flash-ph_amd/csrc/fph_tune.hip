// Autotune table: launch parameters chosen per problem-size bucket.
//
// Four knobs change how work is spread, never what is computed (every value
// gives bit-identical outputs):
//   FPH_TUNE_SCREEN_TILE     points per side of a distance tile of the direct
//                            edge screen (32, 64 or 128), keyed by (n, d);
//   FPH_TUNE_ENUM_BLOCK      threads per workgroup of the link kernels
//                            (triangles, tetrahedra, cofacet columns; one
//                            wavefront per edge or triangle), keyed by the
//                            number of items;
//   FPH_TUNE_APPARENT_BLOCK  columns per workgroup of the apparent-pair and
//                            clearing kernels, keyed by the number of columns;
//   FPH_TUNE_QUEUE_WGS       workgroups of the residual reduction's queue for
//                            one segment (1: one workgroup, ordered on chip;
//                            more: the device-wide queue), keyed by columns.
// A bucket is (floor(log2 size), floor(log2 d)); 0 in a slot means "default".
// Slots are atomics, so lookups from concurrent host threads need no lock.
// The reference tunes Triton block sizes by hand (edge_kernel.py:193-222,
// triangle_kernel.py:366-397); here a host-side tuner times the candidates
// on the caller's data and writes its choice through fph_autotune_set.
#include "fph_common.h"

#include <atomic>

namespace {

constexpr int kBuckets = 40;
std::atomic<int> g_table[FPH_TUNE_NKNOBS][kBuckets][kBuckets];

const int kScreenTiles[] = {32, 64, 128};
const int kBlocks[] = {64, 128, 256, 512, 1024};
const int kQueueWgs[] = {1, 2, 4, 8, 16, 32, 64};

int bucket_of(int64_t x) {
    if (x < 1) return 0;
    int b = 0;
    while (x > 1 && b < kBuckets - 1) {
        x >>= 1;
        b++;
    }
    return b;
}

int default_of(int knob, int64_t size) {
    switch (knob) {
        // 128 won at every (n, d) measured (n 2K-20K, d 4-24: 1.6-2.3x over 64)
        case FPH_TUNE_SCREEN_TILE: return 128;
        case FPH_TUNE_ENUM_BLOCK: return 256;
        case FPH_TUNE_APPARENT_BLOCK: return 256;
        // the device-wide queue pays a few extra launches; from 8192 columns
        // on its 16 workgroups win (n=10K, d=10: 2.4 -> 1.3 ms for H1)
        case FPH_TUNE_QUEUE_WGS: return size >= 8192 ? 16 : 1;
        default: return 0;
    }
}

int candidates_of(int knob, const int** list) {
    if (knob == FPH_TUNE_SCREEN_TILE) {
        *list = kScreenTiles;
        return (int)(sizeof(kScreenTiles) / sizeof(int));
    }
    if (knob == FPH_TUNE_ENUM_BLOCK || knob == FPH_TUNE_APPARENT_BLOCK) {
        *list = kBlocks;
        return (int)(sizeof(kBlocks) / sizeof(int));
    }
    if (knob == FPH_TUNE_QUEUE_WGS) {
        *list = kQueueWgs;
        return (int)(sizeof(kQueueWgs) / sizeof(int));
    }
    return 0;
}

}  // namespace

int fph_tune_value(int knob, int64_t size, int64_t d) {
    if (knob < 0 || knob >= FPH_TUNE_NKNOBS) return 0;
    const int v = g_table[knob][bucket_of(size)][bucket_of(d)].load(std::memory_order_relaxed);
    return v ? v : default_of(knob, size);
}

extern "C" int fph_autotune_candidates(int knob, int32_t* out, int cap) {
    const int* list = nullptr;
    const int m = candidates_of(knob, &list);
    if (m == 0 || cap < 0) return FPH_ERR_ARG;
    for (int i = 0; i < m && i < cap; i++) out[i] = list[i];
    return m;
}

extern "C" int fph_autotune_set(int knob, int64_t size, int64_t d, int value) {
    const int* list = nullptr;
    const int m = candidates_of(knob, &list);
    if (m == 0 || size < 0 || d < 0) return FPH_ERR_ARG;
    bool ok = value == 0;
    for (int i = 0; i < m; i++) ok = ok || list[i] == value;
    if (!ok) return FPH_ERR_ARG;
    g_table[knob][bucket_of(size)][bucket_of(d)].store(value, std::memory_order_relaxed);
    return FPH_OK;
}

extern "C" int fph_autotune_query(int knob, int64_t size, int64_t d) {
    if (knob < 0 || knob >= FPH_TUNE_NKNOBS || size < 0 || d < 0) return FPH_ERR_ARG;
    return fph_tune_value(knob, size, d);
}

extern "C" void fph_autotune_reset(void) {
    for (int k = 0; k < FPH_TUNE_NKNOBS; k++)
        for (int a = 0; a < kBuckets; a++)
            for (int b = 0; b < kBuckets; b++) g_table[k][a][b].store(0, std::memory_order_relaxed);
}
