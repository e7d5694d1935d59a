// gvdbx_reach.h -- extra-brick reach of binned points (device helpers shared
// by the binning kernels).
#pragma once
#include "gvdbx_kernels.h"

namespace {

// Extra-brick reach of every level-0 cell from the cell-sorted points: each
// point's box of gvdbCalcExtraBrickId (cuda_gvdb_particles.cuh:493-494) as a
// 3x3x3 mask; the masks of one cell are OR-ed within the wave (points of a
// cell are contiguous) and one atomicOr per cell run lands in reach0.
// Bits dx+1 for dx in [lo, hi] (-1 <= lo, hi <= 1); 0 when hi < lo.
__device__ __forceinline__ uint32_t reach_range3(int lo, int hi) {
    return hi < lo ? 0u : ((2u << (hi + 1)) - 1u) ^ ((1u << (lo + 1)) - 1u);
}

// Bit (dz+1)*9 + (dy+1)*3 + (dx+1) for every cell offset in the box: the
// per-axis bit sets spread to positions {0,1,2} x {0,3,6} x {0,9,18} and
// multiplied (no two terms share a bit, so no carries).
__device__ __forceinline__ uint32_t reach_mask(const int (&lo)[3], const int (&hi)[3]) {
    const uint32_t bx = reach_range3(lo[0], hi[0]), by = reach_range3(lo[1], hi[1]), bz = reach_range3(lo[2], hi[2]);
    const uint32_t ym = (by & 1u) | ((by & 2u) << 2) | ((by & 4u) << 4);
    const uint32_t zm = (bz & 1u) | ((bz & 2u) << 8) | ((bz & 4u) << 16);
    return bx * ym * zm;
}

constexpr int REACH_UNROLL = 4;

// Home cell (key, -1 if outside g0) and reach mask of one binned point w
// (point index idx for the batched clouds' cell shift).
__device__ __forceinline__ void reach_point(const float (&w)[3], int idx, const ReachParams& rp, int& key, uint32_t& m) {
    key = -1;
    m = 0;
    int c[3], sh[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = gvdbx_cell(w[a], rp.logr0);
    if (rp.cm.ncloud > 0) {
        const int cl = rp.cm.cloud_of_point(idx);
#pragma unroll
        for (int a = 0; a < 3; a++) sh[a] = rp.cm.cell_off[3 * cl + a];
    }
    const int64_t hc = rp.g0.find(c[0] + sh[0], c[1] + sh[1], c[2] + sh[2]);
    if (hc < 0) return;
    key = (int)hc;
    m = 1u << 13;
    if (rp.use_mask) {
        int lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            int ix = (int)w[a];
            int bmn = (int)((float)ix - rp.radius);
            int bmx = (int)((float)ix + 2.0f * rp.radius);
            lo[a] = max(gvdbx_cdiv(bmn, rp.logr0) - c[a], -1);
            hi[a] = min(gvdbx_cdiv(bmx, rp.logr0) - c[a], 1);
        }
        m = reach_mask(lo, hi);
    }
}

// ORs the masks of equal keys over the wave (all 64 lanes take part; lanes
// with consecutive bin entries) and sets each run's bits with one atomicOr.
__device__ __forceinline__ void reach_commit(int key, uint32_t m, uint32_t* reach0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t om = __shfl_down(m, d, 64);
        int ok = __shfl_down(key, d, 64);
        if (lane + d < 64 && ok == key) m |= om;
    }
    int prev = __shfl_up(key, 1, 64);
    if (key >= 0 && (lane == 0 || prev != key)) atomicOr(&reach0[key], m);
}

}  // namespace
