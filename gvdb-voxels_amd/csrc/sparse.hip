// sparse.hip -- sparse cell grids for topologies whose bounding box holds
// more level-0 cells than dense per-cell arrays can cover.
//
// The cells a level needs are found as the reference finds its nodes
// (gvdbCalcBrickId cuda_gvdb_particles.cuh:352-382 keys, RadixSortByByte,
// gvdbFindUnique / gvdbCompactUnique 336-350): 64-bit cell keys are sorted
// with the LSD radix sort, run heads are flagged, and a prefix sum over the
// flags gives every distinct key its rank.  Keys are packed in the level's
// box ((x-ox) << (by+bz)) | ((y-oy) << bz) | (z-oz), which orders them as the
// reference's u16 {z,y,x,lev} keys and lets the sort cover only the bits the
// box needs.  A rank is then a cell's index in every per-cell array of the
// topology (flags, reach, node map, bins), and an open-addressing hash table
// maps a key back to its rank for the lookups (CellGrid::find).
#include "gvdbx_kernels.h"

namespace {

inline unsigned sp_grid(int64_t n, int threads = 256, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// packed home cell of every valid point, `none` for invalid points
__global__ __launch_bounds__(256) void k_sp_point_keys(PointSrc src, int64_t n, float tx, float ty, float tz,
                                                       CellGrid box, int logr0, uint64_t none, uint64_t* keys) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float w[3];
        uint64_t k = none;
        if (gvdbx_point(src, i, tx, ty, tz, w[0], w[1], w[2])) {
            const int cx = gvdbx_cell(w[0], logr0), cy = gvdbx_cell(w[1], logr0), cz = gvdbx_cell(w[2], logr0);
            if (box.in_box(cx, cy, cz)) k = box.pack(cx, cy, cz);
        }
        keys[i] = k;
    }
}

// flag[i] = 1 at the head of each run of equal keys (keys sorted), `none` excluded
__global__ __launch_bounds__(256) void k_sp_heads(const uint64_t* keys, int64_t n, uint64_t none, int32_t* flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        flag[i] = (k != none && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void k_sp_compact(const uint64_t* keys, const int32_t* flag, const int32_t* off,
                                                    int64_t n, uint64_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (flag[i]) out[off[i]] = keys[i];
}

// Every cell at offsets [lo, hi]^3 of each input cell (packed in box gi),
// packed in box go, or `none` where it leaves go: w^3 outputs per input.
__global__ __launch_bounds__(256) void k_sp_dilate(const uint64_t* in, int64_t m, CellGrid gi, int lo, int hi,
                                                   CellGrid go, uint64_t none, uint64_t* out) {
    const int w = hi - lo + 1, w3 = w * w * w;
    const int64_t total = m * w3;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = j / w3;
        const int d = (int)(j - i * w3);
        int x, y, z;
        gi.unpack(in[i], x, y, z);
        x += d / (w * w) + lo;
        y += (d / w) % w + lo;
        z += d % w + lo;
        out[j] = go.in_box(x, y, z) ? go.pack(x, y, z) : none;
    }
}

// parent cell (child >> logres per axis) of every input cell, packed in box gp
__global__ __launch_bounds__(256) void k_sp_parents(const uint64_t* in, int64_t m, CellGrid gc, int logres,
                                                    CellGrid gp, uint64_t none, uint64_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        gc.unpack(in[i], x, y, z);
        x >>= logres;
        y >>= logres;
        z >>= logres;
        out[i] = gp.in_box(x, y, z) ? gp.pack(x, y, z) : none;
    }
}

// insert key i with rank i (keys distinct); the table has more slots than keys
__global__ __launch_bounds__(256) void k_sp_hash(const uint64_t* keys, int64_t m, uint4* ht, uint32_t hmask) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        for (uint32_t s = CellGrid::slot_of(k) & hmask;; s = (s + 1) & hmask) {
            unsigned long long* slot = (unsigned long long*)&ht[s];
            if (atomicCAS(slot, ~0ull, (unsigned long long)k) == ~0ull) {
                ht[s].z = (uint32_t)i;
                break;
            }
        }
    }
}

}  // namespace

void launch_sp_point_keys(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid box, int logr0,
                          uint64_t none, uint64_t* keys, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_sp_point_keys, dim3(sp_grid(n)), dim3(256), 0, s, src, n, tx, ty, tz, box, logr0, none,
                           keys);
}

void launch_sp_heads(const uint64_t* keys, int64_t n, uint64_t none, int32_t* flag, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_sp_heads, dim3(sp_grid(n)), dim3(256), 0, s, keys, n, none, flag);
}

void launch_sp_compact(const uint64_t* keys, const int32_t* flag, const int32_t* off, int64_t n, uint64_t* out,
                       hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_sp_compact, dim3(sp_grid(n)), dim3(256), 0, s, keys, flag, off, n, out);
}

void launch_sp_dilate(const uint64_t* in, int64_t m, CellGrid gi, int lo, int hi, CellGrid go, uint64_t none,
                      uint64_t* out, hipStream_t s) {
    const int w = hi - lo + 1;
    if (m > 0)
        hipLaunchKernelGGL(k_sp_dilate, dim3(sp_grid(m * w * w * w)), dim3(256), 0, s, in, m, gi, lo, hi, go, none,
                           out);
}

void launch_sp_parents(const uint64_t* in, int64_t m, CellGrid gc, int logres, CellGrid gp, uint64_t none,
                       uint64_t* out, hipStream_t s) {
    if (m > 0) hipLaunchKernelGGL(k_sp_parents, dim3(sp_grid(m)), dim3(256), 0, s, in, m, gc, logres, gp, none, out);
}

void launch_sp_hash(const uint64_t* keys, int64_t m, uint4* ht, uint32_t hmask, hipStream_t s) {
    if (m > 0) hipLaunchKernelGGL(k_sp_hash, dim3(sp_grid(m)), dim3(256), 0, s, keys, m, ht, hmask);
}
