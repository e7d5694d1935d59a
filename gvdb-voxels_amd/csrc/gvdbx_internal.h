// gvdbx_internal.h -- shared definitions of the MI355X point -> sparse-volume
// library (host runtime + gfx950 kernels).  Not part of the public C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GVDBX_MAXLEV 10
#define GVDBX_MAXCHAN 8
#define GVDBX_ID_UNDEFL 0xFFFFFFFFu
#define GVDBX_ID_UNDEF64 0xFFFFFFFFFFFFFFFFull
#define GVDBX_NOHIT 1.0e10f

// 64-byte node record; layout of the reference Node (gvdb_node.h:27-40,
// cuda_gvdb_nodes.cuh:24-35) so pools read back byte-identical.
struct alignas(16) GNode {
    uint8_t mLev, mFlags, mPriority, pad;
    int32_t mPos[3];
    int32_t mValue[3];
    float mVRange[3];
    uint64_t mParent;
    uint64_t mChildList;
    uint64_t mMask;
};
static_assert(sizeof(GNode) == 64, "node record must be 64 bytes");

// atlas map entry (cuda_gvdb_nodes.cuh:37-40)
struct alignas(16) GAtlasNode {
    int32_t mPos[3];
    int32_t mLeafID;
};
static_assert(sizeof(GAtlasNode) == 16, "atlas map entry must be 16 bytes");

// A dense cell grid over one tree level: cell (x,y,z) of the level (node
// corner / range) maps to index ((x-ox)*ny + (y-oy))*nz + (z-oz), which is
// the reference's sorted key order (x most significant).
struct CellGrid {
    int ox, oy, oz;
    int nx, ny, nz;
    __host__ __device__ int64_t count() const { return (int64_t)nx * ny * nz; }
    __host__ __device__ bool contains(int x, int y, int z) const {
        return x >= ox && y >= oy && z >= oz && x < ox + nx && y < oy + ny && z < oz + nz;
    }
    __host__ __device__ int64_t index(int x, int y, int z) const {
        return ((int64_t)(x - ox) * ny + (y - oy)) * nz + (z - oz);
    }
};

// Point input as the reference DataPtr handles it: base + offset + stride.
struct PointSrc {
    const char* base;
    int64_t stride;
};

struct BBoxAccum {
    int32_t mn[3];   // float bits of non-negative floats: integer order == float order
    int32_t mx[3];
    unsigned long long nvalid;
};

// Validity rule shared by every stage: finite, z != NOHIT before the
// translation, and no coordinate < 0 after it.
__device__ __forceinline__ bool gvdbx_point(const PointSrc& s, int64_t i, float tx, float ty, float tz,
                                            float& wx, float& wy, float& wz) {
    const float* p = (const float*)(s.base + i * s.stride);
    float x = p[0], y = p[1], z = p[2];
    if (!(__builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z))) return false;
    if (z == GVDBX_NOHIT) return false;
    wx = x + tx;
    wy = y + ty;
    wz = z + tz;
    if (!(__builtin_isfinite(wx) && __builtin_isfinite(wy) && __builtin_isfinite(wz))) return false;
    return wx >= 0.0f && wy >= 0.0f && wz >= 0.0f;
}

// Covering-node corner along one axis as the reference computes it
// (GetCoveringNode, cuda_gvdb_particles.cuh:88-100).
__device__ __forceinline__ int gvdbx_cover(float p, int range) {
    float q = p / (float)range;
    int np = (int)(ceilf(q) * (float)range);
    if (p < (float)np) np -= range;
    return np;
}

// Squared distance in the fixed operation order used everywhere.
__device__ __forceinline__ float gvdbx_d2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

#define GVDBX_CHECK(call)                                                          \
    do {                                                                            \
        hipError_t _e = (call);                                                     \
        if (_e != hipSuccess) {                                                     \
            gvdbx_set_error("HIP error %s at %s:%d", hipGetErrorString(_e), __FILE__, __LINE__); \
            return GVDBX_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

void gvdbx_set_error(const char* fmt, ...);
