// color.hip -- uchar4 colour channels on the points path: the colour mean of
// GatherDensity / GatherLevelSet (cuda_gvdb_particles.cuh:718-731,757-771)
// and the colour part of ScatterDensity (cuda_gvdb_particles.cuh:554-572,
// gvdbScatterPointAvgCol 610-642).
//
// A colour voxel is one 32-bit word, x in the lowest byte (uchar4).  Colour
// sums are exact integers, so no result depends on the order in which
// points are met: the subcell lists and scatter races of the reference, whose
// order is that of atomics, have no order here at all.
#include "gvdbx_kernels.h"

namespace {

// float -> uchar as CUDA's saturating conversion: truncation toward zero,
// clamped to [0, 255], NaN -> 0.
__device__ __forceinline__ uint32_t sat_u8(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 255.0f) return 255u;
    return (uint32_t)x;
}

__device__ __forceinline__ uint32_t load_rgba(const char* base, int64_t stride, int64_t i) {
    return *(const uint32_t*)(base + i * stride);
}

constexpr int CLR_MAX_SUB = 1024;   // subcells accumulated per pass (5 words each in LDS)

// One workgroup per leaf.  The leaf's subcells are taken in passes of at
// most CLR_MAX_SUB; in each pass every point of the 27 bins around the leaf
// (a point's insertion box reaches at most one cell beyond its own when
// radius + subcell <= brick) adds its colour and a count to every subcell
// of the pass inside its insertion box, with LDS integer atomics.  The
// insertion box is InsertPointsSubcell's (cuda_gvdb_particles.cuh:237-242):
// [(int(p - R) / sc) * sc, (int(p + 1 + R) / sc) * sc] per axis, subcell
// origins inclusive.  Then every interior voxel of the pass's subcells gets
//   f = init + (float)S / 255,  out = sat_u8((f / (float)(cnt + accum)) * 255)
// per component, init = (0,0,0,1), or the stored colour / 255 when
// accumulating.
__global__ __launch_bounds__(256) void k_gather_color(ColorGatherParams cp) {
    const int64_t k = blockIdx.x;
    if (k >= cp.nleaf) return;
    __shared__ uint32_t acc[CLR_MAX_SUB * 5];
    const int r0 = cp.res0, sc = cp.sc, nsd = r0 / sc;
    const int nsub = nsd * nsd * nsd;
    const GNode nd = cp.leaves[k];
    const int cx = nd.mPos[0] >> cp.logr0, cy = nd.mPos[1] >> cp.logr0, cz = nd.mPos[2] >> cp.logr0;
    const int br = cp.brick_res, a = cp.apron;
    uint32_t* brick = cp.chan + k * (int64_t)br * br * br;
    for (int s0 = 0; s0 < nsub; s0 += CLR_MAX_SUB) {
        const int s1 = min(nsub, s0 + CLR_MAX_SUB);
        for (int i = threadIdx.x; i < (s1 - s0) * 5; i += blockDim.x) acc[i] = 0u;
        __syncthreads();
        for (int d = 0; d < 27; d++) {
            const int bx = cx + d % 3 - 1, by = cy + (d / 3) % 3 - 1, bz = cz + d / 9 - 1;
            const int64_t b = cp.gb.find(bx, by, bz);
            if (b < 0) continue;
            const int32_t beg = cp.off[b], end = cp.off[b + 1];
            for (int32_t j = beg + (int32_t)threadIdx.x; j < end; j += (int32_t)blockDim.x) {
                const float4 q = cp.sorted[j];
                const float pw[3] = {q.x, q.y, q.z};
                int lo[3], hi[3];
                bool empty = false;
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    // subcell indices of the leaf inside the box
                    const int blo = gvdbx_cdiv_mul((int)(pw[t] - cp.R), cp.sc_shift);
                    const int bhi = gvdbx_cdiv_mul((int)(pw[t] + 1 + cp.R), cp.sc_shift);
                    lo[t] = max(blo - nd.mPos[t], 0) >> cp.sc_shift;
                    const int h = bhi - nd.mPos[t];
                    hi[t] = h < 0 ? -1 : min(h >> cp.sc_shift, nsd - 1);
                    empty = empty || lo[t] > hi[t];
                }
                if (empty) continue;
                const uint32_t c = load_rgba(cp.clr, cp.clr_stride, (int64_t)__float_as_int(q.w));
                for (int sz = lo[2]; sz <= hi[2]; sz++)
                    for (int sy = lo[1]; sy <= hi[1]; sy++)
                        for (int sx = lo[0]; sx <= hi[0]; sx++) {
                            const int s = (sz * nsd + sy) * nsd + sx;
                            if (s < s0 || s >= s1) continue;
                            uint32_t* e = acc + (s - s0) * 5;
                            atomicAdd(e + 0, 1u);
                            atomicAdd(e + 1, c & 255u);
                            atomicAdd(e + 2, (c >> 8) & 255u);
                            atomicAdd(e + 3, (c >> 16) & 255u);
                            atomicAdd(e + 4, c >> 24);
                        }
            }
        }
        __syncthreads();
        // interior voxels of the pass's subcells
        const int nv = (s1 - s0) * sc * sc * sc;
        for (int i = threadIdx.x; i < nv; i += blockDim.x) {
            const int s = s0 + i / (sc * sc * sc), v = i % (sc * sc * sc);
            const int x = (s % nsd) * sc + v % sc, y = ((s / nsd) % nsd) * sc + (v / sc) % sc,
                      z = (s / (nsd * nsd)) * sc + v / (sc * sc);
            const uint32_t* e = acc + (s - s0) * 5;
            const float n = (float)(e[0] + (cp.accum ? 1u : 0u));
            uint32_t* dst = brick + ((int64_t)(z + a) * br + (y + a)) * br + (x + a);
            const uint32_t old = cp.accum ? *dst : 0u;
            uint32_t out = 0u;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float init = cp.accum ? (float)((old >> (8 * t)) & 255u) / 255.0f : (t == 3 ? 1.0f : 0.0f);
                const float f = init + (float)e[1 + t] / 255.0f;
                out |= sat_u8((f / n) * 255.0f) << (8 * t);
            }
            *dst = out;
        }
        __syncthreads();
    }
}

constexpr int CLR_MAX_VOX = 2048;   // voxels accumulated per pass

// One workgroup per leaf over the points of its own bin (the points whose
// home voxel it holds), in passes of CLR_MAX_VOX voxels.  avg: per voxel the
// count and the x, y, z sums (uint32, as the reference's colour buffer),
// then (Sx/cnt, Sy/cnt, Sz/cnt, 255) where cnt > 0; otherwise each voxel
// takes the colour of its highest-numbered point (LDS atomicMax of the
// point index).  Voxels no point lands in keep their value.
__global__ __launch_bounds__(256) void k_scatter_color(ColorScatterParams cp) {
    const int64_t k = blockIdx.x;
    if (k >= cp.nleaf) return;
    __shared__ uint32_t acc[CLR_MAX_VOX * 4];
    const int r0 = cp.res0, nvox = r0 * r0 * r0;
    const GNode nd = cp.leaves[k];
    const int cx = nd.mPos[0] >> cp.logr0, cy = nd.mPos[1] >> cp.logr0, cz = nd.mPos[2] >> cp.logr0;
    const int64_t b = cp.gb.find(cx, cy, cz);
    if (b < 0) return;
    const int32_t beg = cp.off[b], end = cp.off[b + 1];
    if (beg == end) return;
    const int br = cp.brick_res, a = cp.apron;
    uint32_t* brick = cp.chan + k * (int64_t)br * br * br;
    for (int v0 = 0; v0 < nvox; v0 += CLR_MAX_VOX) {
        const int v1 = min(nvox, v0 + CLR_MAX_VOX);
        for (int i = threadIdx.x; i < (v1 - v0) * 4; i += blockDim.x) acc[i] = cp.avg ? 0u : 0xFFFFFFFFu;
        __syncthreads();
        for (int32_t j = beg + (int32_t)threadIdx.x; j < end; j += (int32_t)blockDim.x) {
            const float4 q = cp.sorted[j];
            // home voxel: the reference's pi = int(p - mPos) and its range test
            const float p[3] = {q.x - (float)nd.mPos[0], q.y - (float)nd.mPos[1], q.z - (float)nd.mPos[2]};
            const int pi[3] = {(int)p[0], (int)p[1], (int)p[2]};
            if (pi[0] < 0 || pi[1] < 0 || pi[2] < 0 || pi[0] >= r0 || pi[1] >= r0 || pi[2] >= r0) continue;
            const int vid = (pi[2] * r0 + pi[1]) * r0 + pi[0];
            if (vid < v0 || vid >= v1) continue;
            const int32_t idx = __float_as_int(q.w);
            if (cp.avg) {
                const uint32_t c = load_rgba(cp.clr, cp.clr_stride, idx);
                uint32_t* e = acc + (vid - v0) * 4;
                atomicAdd(e + 0, 1u);
                atomicAdd(e + 1, c & 255u);
                atomicAdd(e + 2, (c >> 8) & 255u);
                atomicAdd(e + 3, (c >> 16) & 255u);
            } else {
                atomicMax((int32_t*)acc + (vid - v0), idx);
            }
        }
        __syncthreads();
        for (int v = v0 + (int)threadIdx.x; v < v1; v += (int)blockDim.x) {
            const int x = v % r0, y = (v / r0) % r0, z = v / (r0 * r0);
            uint32_t* dst = brick + ((int64_t)(z + a) * br + (y + a)) * br + (x + a);
            if (cp.avg) {
                const uint32_t* e = acc + (v - v0) * 4;
                if (e[0] > 0u) *dst = (e[1] / e[0]) | (e[2] / e[0]) << 8 | (e[3] / e[0]) << 16 | 255u << 24;
            } else {
                const int32_t idx = ((const int32_t*)acc)[v - v0];
                if (idx >= 0) *dst = load_rgba(cp.clr, cp.clr_stride, idx);
            }
        }
        __syncthreads();
    }
}

}  // namespace

int color_max_subcells() { return CLR_MAX_SUB; }

void launch_gather_color(ColorGatherParams cp, hipStream_t s) {
    if (cp.nleaf > 0)
        hipLaunchKernelGGL(k_gather_color, dim3((unsigned)cp.nleaf), dim3(256), 0, s, cp);
}

void launch_scatter_color(ColorScatterParams cp, hipStream_t s) {
    if (cp.nleaf > 0)
        hipLaunchKernelGGL(k_scatter_color, dim3((unsigned)cp.nleaf), dim3(256), 0, s, cp);
}
