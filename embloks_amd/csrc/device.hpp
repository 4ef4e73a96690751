// device.hpp - device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "common.hpp"

namespace emb {

// Scene owning global block `blk` of kernel class `kc` (prefix has n_scenes + 1
// entries). All lanes compute the same answer; readfirstlane makes it scalar so
// the scene's fields load through the scalar cache.
__device__ __forceinline__ int scene_of(const Batch &b, int kc, int blk) {
    const int32_t *p = b.prefix + kc * (b.n_scenes + 1);
    int lo = 0, hi = b.n_scenes;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (p[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

__device__ __forceinline__ int block_in_scene(const Batch &b, int kc, int s, int blk) {
    return blk - b.prefix[kc * (b.n_scenes + 1) + s];
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Flat cell of an electron, clamped onto the grid (physics.rs:24-29).
__device__ __forceinline__ int cell_of(const DevScene &S, int32_t cxy) {
    const int c = clampi(cell_x(cxy), 0, S.cols - 1);
    const int r = clampi(cell_y(cxy), 0, S.rows - 1);
    return c * S.rows + r;
}

// Start of cell `cell` in the sorted order right after the scan (before the
// canonical pass writes absolute starts into `count`).
__device__ __forceinline__ int scan_start(const DevScene &S, int cell) {
    const int cells = S.cols * S.rows;
    if (cell >= cells) return S.n;
    return S.lstart[cell] + S.toff_pre[cell / SCAN_TILE];
}

__device__ __forceinline__ unsigned long long as_ull(long long v) { return (unsigned long long)v; }

// Wave-wide reductions (64 lanes).
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace emb
