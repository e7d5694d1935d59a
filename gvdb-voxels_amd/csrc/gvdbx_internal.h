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

// Cells of one tree level (node corner / range) in the box [ox, ox+nx) x
// [oy, oy+ny) x [oz, oz+nz), numbered in the reference's sorted key order
// (x most significant).
//
// Dense form (ht == null): every cell of the box, cell (x,y,z) at index
// ((x-ox)*ny + (y-oy))*nz + (z-oz).
//
// Sparse form (ht != null): only the ncell cells listed in `keys`, sorted,
// each a packed key ((x-ox) << (by+bz)) | ((y-oy) << bz) | (z-oz) with by, bz
// the bits of ny, nz (so packed order is key order); a cell's index is its
// rank in `keys`, found through an open-addressing hash table `ht` of
// hmask+1 slots {key lo, key hi, rank, 0} (empty slots hold key ~0, load
// factor <= 1/2).  Used when the box holds more cells than dense per-cell
// arrays can cover.
struct CellGrid {
    int ox, oy, oz;
    int nx, ny, nz;
    const uint4* ht;
    const uint64_t* keys;
    int64_t ncell;
    uint32_t hmask;
    int by, bz;
    int pad_;
    __host__ __device__ bool sparse() const { return ht != nullptr; }
    __host__ __device__ int64_t count() const { return ht ? ncell : (int64_t)nx * ny * nz; }
    __host__ __device__ bool in_box(int x, int y, int z) const {
        return x >= ox && y >= oy && z >= oz && x < ox + nx && y < oy + ny && z < oz + nz;
    }
    __host__ __device__ uint64_t pack(int x, int y, int z) const {
        return ((uint64_t)(x - ox) << (by + bz)) | ((uint64_t)(y - oy) << bz) | (uint64_t)(z - oz);
    }
    __host__ __device__ static uint32_t slot_of(uint64_t k) {
        return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> 32);
    }
    // index of (x,y,z) in the dense form (the cell must lie in the box)
    __host__ __device__ int64_t dense_index(int x, int y, int z) const {
        return ((int64_t)(x - ox) * ny + (y - oy)) * nz + (z - oz);
    }
    // index of cell (x,y,z), -1 when the grid does not hold it
    __device__ __forceinline__ int64_t find(int x, int y, int z) const {
        if (!ht) {
            // dense grids hold at most 2^28 cells (larger boxes take the sparse
            // form), so 32-bit index arithmetic is exact; one unsigned compare
            // per axis covers both faces of the box
            const uint32_t ux = (uint32_t)x - (uint32_t)ox, uy = (uint32_t)y - (uint32_t)oy, uz = (uint32_t)z - (uint32_t)oz;
            if (ux >= (uint32_t)nx || uy >= (uint32_t)ny || uz >= (uint32_t)nz) return -1;
            return (int64_t)((ux * (uint32_t)ny + uy) * (uint32_t)nz + uz);
        }
        if (!in_box(x, y, z)) return -1;
        const uint64_t k = pack(x, y, z);
        for (uint32_t s = slot_of(k) & hmask;; s = (s + 1) & hmask) {
            const uint4 e = ht[s];
            const uint64_t ek = (uint64_t)e.x | ((uint64_t)e.y << 32);
            if (ek == k) return (int64_t)e.z;
            if (ek == ~0ull) return -1;
        }
    }
    // cell of index i (0 <= i < count())
    __device__ __forceinline__ void cell(int64_t i, int& x, int& y, int& z) const {
        if (!ht) {   // i < 2^28 (dense grids): 32-bit divisions
            const uint32_t u = (uint32_t)i, q = u / (uint32_t)nz;
            z = (int)(u - q * (uint32_t)nz) + oz;
            y = (int)(q % (uint32_t)ny) + oy;
            x = (int)(q / (uint32_t)ny) + ox;
            return;
        }
        unpack(keys[i], x, y, z);
    }
    // cell of a packed key
    __host__ __device__ __forceinline__ void unpack(uint64_t k, int& x, int& y, int& z) const {
        z = (int)(k & ((1ull << bz) - 1)) + oz;
        y = (int)((k >> bz) & ((1ull << by) - 1)) + oy;
        x = (int)(k >> (by + bz)) + ox;
    }
};
static_assert(sizeof(CellGrid) == 64, "CellGrid has no padding (grids are compared with memcmp)");

// Bits needed to hold 0 .. n-1.
__host__ __device__ __forceinline__ int gvdbx_bits_for(int64_t n) {
    int b = 0;
    while (((int64_t)1 << b) < n) b++;
    return b;
}

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

// Cell index floor(p / 2^logr) of a valid (finite, non-negative) coordinate.
// Equal to the reference's ceil-then-correct corner divided by the range;
// the multiply by 2^-logr is exact, so no division is needed.
__device__ __forceinline__ int gvdbx_cell(float p, int logr) {
    return (int)floorf(p * __builtin_ldexpf(1.0f, -logr));
}

// C integer division (truncation toward zero) by 2^s without a divide.
__device__ __forceinline__ int gvdbx_cdiv(int x, int s) { return x >= 0 ? (x >> s) : -((-x) >> s); }

// gvdbx_cdiv(x, s) * 2^s (x rounded toward zero to a multiple of 2^s) with
// masks instead of a 32-bit multiply (a quarter-rate instruction).
__device__ __forceinline__ int gvdbx_cdiv_mul(int x, int s) {
    const int m = (1 << s) - 1;
    return x >= 0 ? (x & ~m) : -((-x) & ~m);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Remapping block b to item (b % 8) * per + b / 8 gives
// every XCD a contiguous run of items (spatially adjacent bricks, which share
// neighbour bins), so the data they share stays in one L2.  Launch with
// gvdbx_xcd_blocks(n) blocks; items >= n return -1.
constexpr int GVDBX_NUM_XCD = 8;
__host__ __device__ __forceinline__ int64_t gvdbx_xcd_blocks(int64_t n) {
    return (n + GVDBX_NUM_XCD - 1) / GVDBX_NUM_XCD * GVDBX_NUM_XCD;
}
__device__ __forceinline__ int64_t gvdbx_xcd_item(int64_t b, int64_t n, int remap = 1) {
    if (!remap) return b < n ? b : -1;
    int64_t per = (n + GVDBX_NUM_XCD - 1) / GVDBX_NUM_XCD;
    int64_t t = (b % GVDBX_NUM_XCD) * per + b / GVDBX_NUM_XCD;
    return t < n ? t : -1;
}

// Squared distance in the fixed operation order used everywhere.
__device__ __forceinline__ float gvdbx_d2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Batched clouds share one packed integer index space.  Cloud c owns points
// [pt_off[c], pt_off[c+1]) in its own FP32 coordinates; its level-0 cells
// are shifted by cell_off[3c..3c+2] in the packed space (an integer shift:
// float positions and distances stay cloud-local and exact).  ncloud == 0
// means one cloud and no shift.
struct CloudMap {
    int ncloud;
    const int64_t* pt_off;
    const int32_t* cell_off;
    __device__ __forceinline__ int cloud_of_point(int64_t i) const {
        int lo = 0, hi = ncloud - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pt_off[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }
};

// FP16 subcell lists (InsertPointsSubcell_FP16, cuda_gvdb_particles.cuh:
// 154-216, read back by gvdbGatherLevelSet_fp16 775-811): a position w is
// stored as the ushort (w - mn) / rg * 65535 (truncated; clamped to
// [0, 65535], NaN -> 0) and read back as q / 65535 * rg + mn, every step an
// FP32 operation in that order.  mn, rg: the bounding box GetBoundingBox
// left (pos_min and pos_max - pos_min).
struct Quant16 {
    int on;
    float mn[3], rg[3];
};
__host__ __device__ __forceinline__ float quant16_round_trip(float w, float mn, float rg) {
    const float t = ((w - mn) / rg) * 65535.0f;
    const uint32_t q = !(t > 0.0f) ? 0u : (t >= 65535.0f ? 65535u : (uint32_t)t);
    return ((float)q / 65535.0f) * rg + mn;
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
