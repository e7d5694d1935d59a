// topology.hip -- gfx950 kernels that rebuild the GVDB tree from points.
//
// Design: instead of sorting N*depth 64-bit keys (the reference's Thrust
// path, VG.cpp:904-928), every level of the tree is a dense cell grid over
// the point extents.  One atomicOr per point records the cell it falls in
// plus a 27-bit mask of the neighbouring cells its extra-brick box reaches
// (gvdbCalcExtraBrickId, cuda_gvdb_particles.cuh:467-512).  Ordered scans
// over the grids (x-major, then y, then z == the reference key order
// lev<<48|x<<32|y<<16|z) give each node its pool index directly: base nodes
// first, extra nodes appended, exactly the order ActivateBricksGPU +
// ActivateExtraBricksGPU produce (VG.cpp:1288-1331, 966-988).
#include "gvdbx_kernels.h"
#include "gvdbx_reach.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ U2 operator+(U2 x, U2 y) { return U2{x.a + y.a, x.b + y.b}; }
__device__ __forceinline__ U2 operator-(U2 x, U2 y) { return U2{x.a - y.a, x.b - y.b}; }

__device__ __forceinline__ uint32_t shfl_up_t(uint32_t v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ U2 shfl_up_t(U2 v, int d) { return U2{__shfl_up(v.a, d, 64), __shfl_up(v.b, d, 64)}; }

template <class T>
__device__ __forceinline__ T zero_t();
template <>
__device__ __forceinline__ uint32_t zero_t<uint32_t>() { return 0u; }
template <>
__device__ __forceinline__ U2 zero_t<U2>() { return U2{0u, 0u}; }

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = shfl_up_t(v, d);
        if (lane >= d) v = v + o;
    }
    return v;
}

// exclusive scan of one value per thread across the block; NW = waves per block
template <class T, int NW>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T& total) {
    T inc = wave_incl_scan(v);
    int w = threadIdx.x >> 6;
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    T wofs = zero_t<T>(), tot = zero_t<T>();
#pragma unroll
    for (int i = 0; i < NW; i++) {
        if (i < w) wofs = wofs + lds[i];
        tot = tot + lds[i];
    }
    __syncthreads();
    total = tot;
    return wofs + inc - v;
}

// ---- scan over tiles: phase 2 (single block) ----
template <class T>
__global__ __launch_bounds__(1024) void k_scan_partials(T* partials, int64_t ntiles, T* total) {
    __shared__ T lds[16];
    int64_t per = (ntiles + 1023) / 1024;
    int64_t b = (int64_t)threadIdx.x * per;
    T s = zero_t<T>();
    for (int64_t i = 0; i < per; i++)
        if (b + i < ntiles) s = s + partials[b + i];
    T tot;
    T ex = block_excl_scan<T, 16>(s, lds, tot);
    T run = ex;
    for (int64_t i = 0; i < per; i++)
        if (b + i < ntiles) {
            T x = partials[b + i];
            partials[b + i] = run;
            run = run + x;
        }
    if (threadIdx.x == 0 && total) *total = tot;
}

// ---- flag scan (base, extra) ----
__device__ __forceinline__ U2 flag_val(const uint8_t* base, const uint8_t* all, int64_t i, int64_t n) {
    if (i >= n) return U2{0, 0};
    uint32_t b = base[i] ? 1u : 0u;
    uint32_t e = (!b && all[i]) ? 1u : 0u;
    return U2{b, e};
}

__global__ __launch_bounds__(SCAN_THREADS) void k_flag_reduce(const uint8_t* base, const uint8_t* all, int64_t n,
                                                                U2* partials) {
    __shared__ U2 lds[4];
    int64_t t0 = (int64_t)blockIdx.x * SCAN_TILE;
    U2 s{0, 0};
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s = s + flag_val(base, all, t0 + k * SCAN_THREADS + threadIdx.x, n);
    U2 tot;
    block_excl_scan<U2, 4>(s, lds, tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Every level in one launch: blocks [tile_start[i], tile_start[i+1]) are the
// tiles of level i.
__global__ __launch_bounds__(SCAN_THREADS) void k_flag_apply(FlagApplyBatch fb) {
    int lv = 0;
    while ((int64_t)blockIdx.x >= fb.tile_start[lv + 1]) lv++;
    const uint8_t* base = fb.lv[lv].base;
    const uint8_t* all = fb.lv[lv].all;
    const int64_t n = fb.lv[lv].n;
    const U2* partials = fb.lv[lv].partials;
    const U2* totals = fb.lv[lv].totals;
    const NodeBuildParams& np = fb.lv[lv].np;
    const int64_t tile = (int64_t)blockIdx.x - fb.tile_start[lv];
    __shared__ U2 lds[4];
    int64_t t0 = tile * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    U2 v[SCAN_ITEMS];
    U2 s{0, 0};
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = flag_val(base, all, t0 + k, n);
        s = s + v[k];
    }
    U2 tot;
    U2 run = block_excl_scan<U2, 4>(s, lds, tot) + partials[tile];
    uint32_t nbase = totals->a;
    const CellGrid& g = np.g;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = t0 + k;
        if (i < n) {
            int32_t node = -1;
            if (v[k].a) node = (int32_t)(np.node_off + run.a);
            else if (v[k].b) node = (int32_t)(np.node_off + nbase + run.b);
            if (!np.keep_idx || node >= 0) np.idx[i] = node;
            if (node >= 0) {
                int cx, cy, cz;
                g.cell(i, cx, cy, cz);
                GNode nd;
                nd.mLev = (uint8_t)np.lev;
                nd.mFlags = 1;
                nd.mPriority = 0;
                nd.pad = 0;
                nd.mPos[0] = (int32_t)(cx * np.range);
                nd.mPos[1] = (int32_t)(cy * np.range);
                nd.mPos[2] = (int32_t)(cz * np.range);
                nd.mValue[0] = nd.mValue[1] = nd.mValue[2] = -1;
                nd.mVRange[0] = nd.mVRange[1] = nd.mVRange[2] = 0.0f;
                nd.mParent = GVDBX_ID_UNDEFL;
                nd.mChildList = np.is_leaf ? (uint64_t)GVDBX_ID_UNDEFL
                                           : (1ull | ((uint64_t)np.lev << 8) | ((uint64_t)node << 16));
                nd.mMask = 0;
                np.nodes[node] = nd;
            }
        }
        run = run + v[k];
    }
}

// ---- count scan (int32 counts -> int32 exclusive offsets, off[n] = total) ----
__global__ __launch_bounds__(SCAN_THREADS) void k_count_reduce(const int32_t* cnt, int64_t n, uint32_t* partials) {
    __shared__ uint32_t lds[4];
    int64_t t0 = (int64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = t0 + k * SCAN_THREADS + threadIdx.x;
        if (i < n) s += (uint32_t)cnt[i];
    }
    uint32_t tot;
    block_excl_scan<uint32_t, 4>(s, lds, tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Each thread scans SCAN_ITEMS consecutive counts.  VEC (both arrays 16-byte
// aligned): whole groups of four move as one 16-byte load or store, the
// groups at the end element by element.
template <bool VEC>
__global__ __launch_bounds__(SCAN_THREADS) void k_count_apply(const int32_t* cnt, int64_t n,
                                                                const uint32_t* partials, int32_t* off) {
    __shared__ uint32_t lds[4];
    int64_t t0 = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int g = 0; g < SCAN_ITEMS; g += 4) {
        const int64_t i = t0 + g;
        if (VEC && i + 3 < n) {
            const int4 x = *reinterpret_cast<const int4*>(cnt + i);
            v[g] = (uint32_t)x.x, v[g + 1] = (uint32_t)x.y, v[g + 2] = (uint32_t)x.z, v[g + 3] = (uint32_t)x.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) v[g + k] = i + k < n ? (uint32_t)cnt[i + k] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) s += v[g + k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<uint32_t, 4>(s, lds, tot) + partials[blockIdx.x];
#pragma unroll
    for (int g = 0; g < SCAN_ITEMS; g += 4) {
        const int64_t i = t0 + g;
        int32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o[k] = (int32_t)run;
            run += v[g + k];
        }
        if (VEC && i + 3 < n) {
            *reinterpret_cast<int4*>(off + i) = make_int4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (i + k <= n) off[i + k] = o[k];   // off[n] = the total
        }
    }
}

// ---- bounding box ----
__global__ void k_bbox_init(BBoxAccum* acc) {
    if (threadIdx.x < 3) {
        acc->mn[threadIdx.x] = 0x7F800000;   // +inf bits
        acc->mx[threadIdx.x] = -1;
    }
    if (threadIdx.x == 0) acc->nvalid = 0;
}

// Bounding box of the valid points in two passes: each block reduces a
// strided share (BBOX_UNROLL independent point loads in flight per thread)
// to one partial, and one block folds the partials into acc with the
// integer order of the float bits, as same-address atomics would (no
// atomics: thousands of them on one line serialise in L2).
constexpr int BBOX_UNROLL = 4;
constexpr int BBOX_BLOCKS = 2048;

__device__ __forceinline__ void bbox_add(const PointSrc& src, int64_t i, int64_t n, float tx, float ty, float tz,
                                         float (&mn)[3], float (&mx)[3], unsigned long long& cnt) {
    float x, y, z;
    if (i < n && gvdbx_point(src, i, tx, ty, tz, x, y, z)) {
        cnt++;
        mn[0] = fminf(mn[0], x); mn[1] = fminf(mn[1], y); mn[2] = fminf(mn[2], z);
        mx[0] = fmaxf(mx[0], x); mx[1] = fmaxf(mx[1], y); mx[2] = fmaxf(mx[2], z);
    }
}

__global__ __launch_bounds__(256) void k_bbox(PointSrc src, int64_t n, float tx, float ty, float tz,
                                              BBoxAccum* part) {
    float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    unsigned long long cnt = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += BBOX_UNROLL * stride) {
#pragma unroll
        for (int u = 0; u < BBOX_UNROLL; u++) bbox_add(src, i + u * stride, n, tx, ty, tz, mn, mx, cnt);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], d, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], d, 64));
        }
        cnt += __shfl_xor(cnt, d, 64);
    }
    __shared__ float red[4][6];
    __shared__ unsigned long long redc[4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        for (int a = 0; a < 3; a++) red[w][a] = mn[a], red[w][3 + a] = mx[a];
        redc[w] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            for (int a = 0; a < 3; a++) mn[a] = fminf(mn[a], red[i][a]), mx[a] = fmaxf(mx[a], red[i][3 + a]);
            cnt += redc[i];
        }
        BBoxAccum b;
        for (int a = 0; a < 3; a++) {
            b.mn[a] = cnt ? __float_as_int(mn[a]) : 0x7F800000;
            b.mx[a] = cnt ? __float_as_int(mx[a]) : -1;
        }
        b.nvalid = cnt;
        part[blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(256) void k_bbox_final(const BBoxAccum* part, int nparts, BBoxAccum* acc) {
    int mn[3] = {0x7F800000, 0x7F800000, 0x7F800000}, mx[3] = {-1, -1, -1};
    unsigned long long cnt = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        const BBoxAccum b = part[i];
        for (int a = 0; a < 3; a++) mn[a] = min(mn[a], b.mn[a]), mx[a] = max(mx[a], b.mx[a]);
        cnt += b.nvalid;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = min(mn[a], __shfl_xor(mn[a], d, 64));
            mx[a] = max(mx[a], __shfl_xor(mx[a], d, 64));
        }
        cnt += __shfl_xor(cnt, d, 64);
    }
    __shared__ int red[4][6];
    __shared__ unsigned long long redc[4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        for (int a = 0; a < 3; a++) red[w][a] = mn[a], red[w][3 + a] = mx[a];
        redc[w] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            for (int a = 0; a < 3; a++) mn[a] = min(mn[a], red[i][a]), mx[a] = max(mx[a], red[i][3 + a]);
            cnt += redc[i];
        }
        for (int a = 0; a < 3; a++) acc->mn[a] = mn[a], acc->mx[a] = mx[a];
        acc->nvalid = cnt;
    }
}

// ---- occupancy + reach ----
__global__ __launch_bounds__(256) void k_mark_boxes(PointSrc src, int64_t n, float tx, float ty, float tz,
                                                    CellGrid g0, int r0, float radius, uint8_t* all0) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float w[3];
        if (!gvdbx_point(src, i, tx, ty, tz, w[0], w[1], w[2])) continue;
        int lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            int ix = (int)w[a];
            int bmn = (int)((float)ix - radius);
            int bmx = (int)((float)ix + 2.0f * radius);
            lo[a] = bmn / r0;
            hi[a] = bmx / r0;
        }
        for (int x = lo[0]; x <= hi[0]; x++)
            for (int y = lo[1]; y <= hi[1]; y++)
                for (int z = lo[2]; z <= hi[2]; z++)
                    if (const int64_t c = g0.find(x, y, z); c >= 0) all0[c] = 1;
    }
}

__global__ __launch_bounds__(256) void k_expand_reach(CellGrid g0, const uint32_t* reach0, uint8_t* base0,
                                                      uint8_t* all0, int generic, const int32_t* allowed) {
    int64_t n = g0.count();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = reach0[i];
        base0[i] = m ? 1 : 0;
        if (!m) continue;
        all0[i] = 1;
        if (generic) continue;
        int cx, cy, cz;
        g0.cell(i, cx, cy, cz);
        m &= ~(1u << 13);
        while (m) {
            int b = __ffs(m) - 1;
            m &= m - 1;
            int dx = b % 3 - 1, dy = (b / 3) % 3 - 1, dz = b / 9 - 1;
            const int64_t t = g0.find(cx + dx, cy + dy, cz + dz);
            if (t >= 0 && (!allowed || allowed[t] >= 0)) all0[t] = 1;
        }
    }
}

// Level-l flags from level l-1: every flagged child cell marks its parent
// (the parents' flags start cleared; concurrent stores write the same 1).
// One thread per child cell, so no thread loops over a parent's children.
__global__ __launch_bounds__(256) void k_level_up(CellGrid gc, const uint8_t* base_c, const uint8_t* all_c,
                                                  CellGrid gp, int logres, uint8_t* base_p, uint8_t* all_p) {
    const int64_t n = gc.count();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t b = base_c[i], a = all_c[i];
        if (!(a | b)) continue;
        int x, y, z;
        gc.cell(i, x, y, z);
        const int64_t p = gp.find(x >> logres, y >> logres, z >> logres);
        if (p < 0) continue;
        if (b) base_p[p] = 1;
        all_p[p] = 1;
    }
}

// ---- position-sensitive checksum of 32-bit words (two 64-bit sums) ----
__global__ __launch_bounds__(256) void k_checksum(const uint32_t* p, int64_t n, unsigned long long* out) {
    unsigned long long a = 0, b = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long w = p[i];
        a += w;
        b += w * (unsigned long long)(i % 65521 + 1);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(out, a);
        atomicAdd(out + 1, b);
    }
}

// ---- fills ----
template <class T>
__global__ __launch_bounds__(256) void k_fill(T* p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---- link (gvdbLinkBricks cuda_gvdb_nodes.cuh:470-505 for levels below top_lev) ----
__device__ __forceinline__ void link_one(const LinkParams& lp, int64_t k) {
    const GNode& c = lp.child[k];
    if (lp.need_flag && !c.mFlags) return;
    int64_t p = 0;
    if (!lp.parent_is_root) {
        int pc[3];
#pragma unroll
        for (int a = 0; a < 3; a++) pc[a] = (int)(c.mPos[a] / lp.range_p);
        const int64_t pcell = lp.gp.find(pc[0], pc[1], pc[2]);
        if (pcell < 0) return;
        p = lp.idx_p[pcell];
        if (p < 0 || p >= lp.nparent) return;
    }
    const GNode& pn = lp.parent_nodes[p];
    int64_t pin[3];
#pragma unroll
    for (int a = 0; a < 3; a++) pin[a] = ((int64_t)(c.mPos[a] - pn.mPos[a]) * lp.res_p) / lp.range_p;
    if (pin[0] < 0 || pin[1] < 0 || pin[2] < 0 || pin[0] >= lp.res_p || pin[1] >= lp.res_p || pin[2] >= lp.res_p) return;
    int64_t bit = (pin[2] * lp.res_p + pin[1]) * lp.res_p + pin[0];
    lp.childw[k].mParent = ((uint64_t)p << 16) | ((uint64_t)(lp.lev + 1) << 8);
    uint64_t cndx = pn.mChildList >> 16;
    int64_t len = (int64_t)lp.res_p * lp.res_p * lp.res_p;
    lp.parent_clist[(int64_t)cndx * len + bit] = ((uint64_t)k << 16) | ((uint64_t)lp.lev << 8);
}

// Every level's links in one launch (levels write disjoint lists and parents).
__global__ __launch_bounds__(256) void k_link(LinkBatch lb) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < lb.start[lb.n]; g += (int64_t)gridDim.x * blockDim.x) {
        int i = 0;
        while (g >= lb.start[i + 1]) i++;
        link_one(lb.lv[i], g - lb.start[i]);
    }
}

// Child lists of levels 1.. set to "no child" and the root node written, in one launch.
__global__ __launch_bounds__(256) void k_clist_init(ClistInit ci) {
    if (ci.root_dst && blockIdx.x == 0 && threadIdx.x == 0) *ci.root_dst = ci.root;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ci.start[ci.n]; g += (int64_t)gridDim.x * blockDim.x) {
        int i = 0;
        while (g >= ci.start[i + 1]) i++;
        ci.list[i][g - ci.start[i]] = GVDBX_ID_UNDEF64;
    }
}

// Leaf bounds accumulator: min (3) and max (3) set to their identities.
__global__ void k_bounds_init(int32_t* out6) {
    if (threadIdx.x < 6) out6[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : INT32_MIN;
}

// ---- atlas (UpdateAtlas VG.cpp:2630-2704, getAtlasPos AL.cpp:705-715) ----
__global__ __launch_bounds__(256) void k_atlas_assign(GNode* leaves, int64_t nleaf, int sx, int sy, int br, int apron,
                                                      GAtlasNode* amap, int64_t nmap) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nmap; k += (int64_t)gridDim.x * blockDim.x) {
        GAtlasNode an;
        if (k < nleaf) {
            int64_t a2 = (int64_t)sx * sy;
            int64_t z = k / a2, rem = k - z * a2;
            int64_t y = rem / sx, x = rem - y * sx;
            GNode& nd = leaves[k];
            nd.mValue[0] = (int32_t)(x * br + apron);
            nd.mValue[1] = (int32_t)(y * br + apron);
            nd.mValue[2] = (int32_t)(z * br + apron);
            an.mPos[0] = nd.mPos[0]; an.mPos[1] = nd.mPos[1]; an.mPos[2] = nd.mPos[2];
            an.mLeafID = (int32_t)k;
        } else {
            an.mPos[0] = an.mPos[1] = an.mPos[2] = (int32_t)GVDBX_ID_UNDEFL;
            an.mLeafID = (int32_t)GVDBX_ID_UNDEFL;
        }
        amap[k] = an;
    }
}

__global__ __launch_bounds__(256) void k_leaf_bounds(const GNode* leaves, int64_t nleaf, int range0, int32_t* out6) {
    int v[6] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nleaf; k += (int64_t)gridDim.x * blockDim.x) {
        if (k == 0) {
            // ComputeBounds seeds both corners with leaf 0's position, active or not
#pragma unroll
            for (int a = 0; a < 3; a++) {
                v[a] = min(v[a], leaves[0].mPos[a]);
                v[3 + a] = max(v[3 + a], leaves[0].mPos[a]);
            }
        }
        if (!leaves[k].mFlags) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = min(v[a], leaves[k].mPos[a]);
            v[3 + a] = max(v[3 + a], leaves[k].mPos[a] + range0);
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = min(v[a], __shfl_xor(v[a], d, 64));
            v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], d, 64));
        }
    __shared__ int red[4][6];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int a = 0; a < 6; a++) red[w][a] = v[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++)
            for (int a = 0; a < 3; a++) {
                v[a] = min(v[a], red[i][a]);
                v[3 + a] = max(v[3 + a], red[i][3 + a]);
            }
        for (int a = 0; a < 3; a++) {
            atomicMin(&out6[a], v[a]);
            atomicMax(&out6[3 + a], v[3 + a]);
        }
    }
}

// ---- point binning ----
// Each wave takes 64 consecutive sorted points REACH_UNROLL times per step
// (loads issued together), ORs the masks of equal home cells across the
// wave and sets each cell's bits with one atomicOr.
__global__ __launch_bounds__(256) void k_reach_sorted(const float4* sorted, const int32_t* total_ptr, int64_t nmax,
                                                      ReachParams rp, uint32_t* reach0) {
    const int64_t total = min((int64_t)*total_ptr, nmax);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total; base += REACH_UNROLL * stride) {
        float4 pts[REACH_UNROLL];
#pragma unroll
        for (int u = 0; u < REACH_UNROLL; u++) {
            const int64_t j = base + u * stride + threadIdx.x;
            if (j < total) pts[u] = sorted[j];
        }
#pragma unroll 1
        for (int u = 0; u < REACH_UNROLL; u++) {
            if (base + u * stride >= total) break;
            const int64_t j = base + u * stride + threadIdx.x;
            int key = -1;
            uint32_t m = 0;
            if (j < total) {
                const float w[3] = {pts[u].x, pts[u].y, pts[u].z};
                reach_point(w, __float_as_int(pts[u].w), rp, key, m);
            }
            reach_commit(key, m, reach0);
        }
    }
}

// Subcell list length of InsertPointsSubcell (gvdbCountSubcell,
// cuda_gvdb_particles.cuh:275-315): per point, the subcells of its box that
// lie in a leaf; summed per block, one atomic per block.  The subcell size
// divides the brick, so along each axis the box's subcells (those >= 0) fall
// into at most a few level-0 cells; the count is the sum over the box's
// cells that are leaves of the product of their per-axis subcell counts (one
// leaf test per cell instead of per subcell).
__device__ __forceinline__ int subcells_in_cell(int lo, int hi, int c, int r0, int sc, int sc_shift) {
    // subcells x = lo, lo+sc, ..., hi (multiples of sc = 2^sc_shift) with x >= 0 inside cell c
    const int a = max(max(lo, 0), c * r0), b = min(hi, c * r0 + r0 - sc);
    return b >= a ? ((b - a) >> sc_shift) + 1 : 0;
}

// Subcell list entries of one point.  When the box's cells lie within one
// cell of the point's home cell (the usual case: radius + subcell size within
// a brick), the leaf tests come from the home cell's 27-bit neighbour-leaf
// mask (one load, shared by the cell's points); otherwise each cell of the
// box is tested (bit 13 of its own mask: the cell itself).
__device__ __forceinline__ unsigned long long subcell_count(const float4& p, const BinParams& bp, int r0, int sc) {
    const float w[3] = {p.x, p.y, p.z};
    // In subcell units the box's subcells along axis a are L[a] .. H1[a] - 1
    // (subcells >= 0 only; empty when H1 <= L) and the home cell's first
    // subcell is Q[a]: the ends are those of the box [cdiv(w - r) * sc,
    // cdiv(w + 1 + r) * sc] divided by sc, exactly (both are multiples of sc).
    const int s = bp.sc_shift, ks = bp.logr0 - s, q = 1 << ks;
    int L[3], H1[3], Q[3], hc[3];
    bool near = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        L[a] = max((int)(w[a] - bp.radius), 0) >> s;
        H1[a] = gvdbx_cdiv((int)(w[a] + 1 + bp.radius), s) + 1;
        hc[a] = gvdbx_cell(w[a], bp.logr0);
        Q[a] = hc[a] * q;
        near = near && L[a] >= Q[a] - q && H1[a] <= Q[a] + 2 * q;
    }
    const int64_t hcell = near ? bp.g0.find(hc[0], hc[1], hc[2]) : -1;
    if (hcell >= 0) {
        const uint32_t m = bp.nbmask[hcell];
        // n[a][d]: the box's subcells along axis a inside cell hc[a] + d - 1
        // (its overlap with subcells [Q + (d-1) q, Q + d q); the box lies in
        // these three cells, so the middle one is the rest); at most
        // q = r0 / sc <= 32 each (bricks are at most 32^3), so they pack
        // into bytes and every 24-bit product below stays exact
        int n[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int n0 = max(min(H1[a], Q[a]) - L[a], 0), n2 = max(H1[a] - max(L[a], Q[a] + q), 0);
            n[a][0] = n0;
            n[a][2] = n2;
            n[a][1] = max(H1[a] - L[a], 0) - n0 - n2;
        }
        const uint32_t nx = (uint32_t)n[0][0] | (uint32_t)n[0][1] << 8 | (uint32_t)n[0][2] << 16;
        // a row's three leaf bits spread to bytes: b * 0x4081 puts bit k at
        // bit 8k (no carries for b < 8)
        auto row_bits = [&](int dy, int dz) {
            return __umul24((m >> (dz * 9 + dy * 3)) & 7u, 0x4081u) & 0x010101u;
        };
        uint32_t c32 = 0;
        if (q <= 15) {
            // x counts times the y count as bytes (<= 225: no carries), one
            // dot product per row, one multiply per z layer
            uint32_t P[3];
#pragma unroll
            for (int dy = 0; dy < 3; dy++) P[dy] = __umul24(nx, (uint32_t)n[1][dy]);
#pragma unroll
            for (int dz = 0; dz < 3; dz++) {
                uint32_t sz = 0;
#pragma unroll
                for (int dy = 0; dy < 3; dy++) sz = __builtin_amdgcn_udot4(P[dy], row_bits(dy, dz), sz, false);
                c32 += __umul24(sz, (uint32_t)n[2][dz]);
            }
        } else {
#pragma unroll
            for (int dz = 0; dz < 3; dz++)
#pragma unroll
                for (int dy = 0; dy < 3; dy++) {
                    const uint32_t row = __builtin_amdgcn_udot4(nx, row_bits(dy, dz), 0u, false);
                    c32 += __umul24(row, __umul24(n[1][dy], n[2][dz]));
                }
        }
        return c32;
    }
    unsigned long long cnt = 0;
    int lo[3], hi[3], c0[3], c1[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = gvdbx_cdiv_mul((int)(w[a] - bp.radius), s);
        hi[a] = gvdbx_cdiv_mul((int)(w[a] + 1 + bp.radius), s);
        c0[a] = max(lo[a], 0) >> bp.logr0;
        c1[a] = hi[a] >> bp.logr0;   // hi < 0: empty range (c1 < c0 or no subcell counted)
    }
    for (int cx = c0[0]; cx <= c1[0]; cx++) {
        const int nx = subcells_in_cell(lo[0], hi[0], cx, r0, sc, bp.sc_shift);
        if (nx == 0) continue;
        for (int cy = c0[1]; cy <= c1[1]; cy++) {
            const int ny = subcells_in_cell(lo[1], hi[1], cy, r0, sc, bp.sc_shift);
            if (ny == 0) continue;
            for (int cz = c0[2]; cz <= c1[2]; cz++) {
                const int nz = subcells_in_cell(lo[2], hi[2], cz, r0, sc, bp.sc_shift);
                if (nz == 0) continue;
                const int64_t ci = bp.g0.find(cx, cy, cz);
                if (ci < 0) continue;
                if ((bp.nbmask[ci] >> 13) & 1u) cnt += (unsigned long long)(nx * ny * nz);   // the cell itself
            }
        }
    }
    return cnt;
}

// SUBCELL_UNROLL points per thread per step, their loads issued together.
constexpr int SUBCELL_UNROLL = 4;

__global__ __launch_bounds__(256) void k_subcell_len(const float4* sorted, const int32_t* total_ptr, int64_t nmax,
                                                     BinParams bp, unsigned long long* len) {
    const int64_t total = min((int64_t)*total_ptr, nmax);
    const int r0 = 1 << bp.logr0, sc = bp.sc;
    unsigned long long lsum = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j0 < total; j0 += SUBCELL_UNROLL * stride) {
        float4 pts[SUBCELL_UNROLL];
#pragma unroll
        for (int u = 0; u < SUBCELL_UNROLL; u++) {
            const int64_t j = j0 + u * stride;
            if (j < total) pts[u] = sorted[j];
        }
#pragma unroll 1
        for (int u = 0; u < SUBCELL_UNROLL && j0 + u * stride < total; u++) lsum += subcell_count(pts[u], bp, r0, sc);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) lsum += __shfl_xor(lsum, d, 64);
    __shared__ unsigned long long red[4];
    if (lane_id() == 0) red[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) lsum += red[i];
        if (lsum) atomicAdd(len, lsum);
    }
}

// Binning by sort: the key of a point is its home cell in the bin grid
// (ncell for points that are invalid or outside it, so they sort last), the
// value (optional) its index.  A stable sort then leaves every bin in
// ascending point order.
__global__ __launch_bounds__(256) void k_bin_keys(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid gb,
                                                  int logr0, CloudMap cm, uint32_t* key, uint32_t* val) {
    const uint32_t none = (uint32_t)gb.count();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float w[3];
        uint32_t k = none;
        if (gvdbx_point(src, i, tx, ty, tz, w[0], w[1], w[2])) {
            int c[3];
#pragma unroll
            for (int a = 0; a < 3; a++) c[a] = gvdbx_cell(w[a], logr0);
            if (cm.ncloud > 0) {
                const int cl = cm.cloud_of_point(i);
#pragma unroll
                for (int a = 0; a < 3; a++) c[a] += cm.cell_off[3 * cl + a];
            }
            const int64_t b = gb.find(c[0], c[1], c[2]);
            if (b >= 0) k = (uint32_t)b;
        }
        key[i] = k;
        if (val) val[i] = (uint32_t)i;
    }
}

// BIN_UNROLL entries per thread per step: their keys and indices are loaded
// together, then all their (random) point reads are in flight at once.
constexpr int BIN_UNROLL = 4;

__global__ __launch_bounds__(256) void k_bin_gather(PointSrc src, int64_t n, float tx, float ty, float tz,
                                                    const uint32_t* key, const uint32_t* val, uint32_t ncell,
                                                    float4* sorted, ReachParams rp, uint32_t* reach0) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // wave-uniform steps, so that the reach pass below can shuffle across the wave
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < n; b0 += BIN_UNROLL * stride) {
        const int64_t j0 = b0 + threadIdx.x;
        uint32_t kk[BIN_UNROLL], ii[BIN_UNROLL];
#pragma unroll
        for (int u = 0; u < BIN_UNROLL; u++) {
            const int64_t j = j0 + u * stride;
            kk[u] = j < n ? key[j] : ncell;
            ii[u] = j < n ? val[j] : 0u;
        }
        float w[BIN_UNROLL][3];
#pragma unroll
        for (int u = 0; u < BIN_UNROLL; u++)
            if (kk[u] < ncell) gvdbx_point(src, ii[u], tx, ty, tz, w[u][0], w[u][1], w[u][2]);
#pragma unroll
        for (int u = 0; u < BIN_UNROLL; u++)
            if (kk[u] < ncell) sorted[j0 + u * stride] = make_float4(w[u][0], w[u][1], w[u][2], __int_as_float((int)ii[u]));
        if (reach0) {
            // the extra-brick reach of these points while they are in registers
            // (the work of k_reach_sorted: the valid entries are the binned ones)
#pragma unroll 1
            for (int u = 0; u < BIN_UNROLL; u++) {
                if (b0 + u * stride >= n) break;
                int rk = -1;
                uint32_t m = 0;
                if (kk[u] < ncell) {
                    const float wu[3] = {w[u][0], w[u][1], w[u][2]};
                    reach_point(wu, (int)ii[u], rp, rk, m);
                }
                reach_commit(rk, m, reach0);
            }
        }
    }
}

// off[c] = lower bound of c in the sorted keys (binary search per cell)
__global__ __launch_bounds__(256) void k_bin_offsets(const uint32_t* key, int64_t n, int64_t ncell, int32_t* off) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= ncell; c += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)key[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        off[c] = (int32_t)lo;
    }
}

// Per level-0 cell, bit (dz+1)*9 + (dy+1)*3 + (dx+1) set when the cell at
// offset (dx, dy, dz) is a leaf.
__global__ __launch_bounds__(256) void k_leaf_nbmask(const int32_t* idx0, CellGrid g0, uint32_t* mask,
                                                     unsigned long long* reset) {
    const int64_t n = g0.count();
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0ull;   // a counter the next pass adds into
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        g0.cell(c, x, y, z);
        uint32_t m = 0;
        for (int d = 0; d < 27; d++) {
            const int nx = x + d % 3 - 1, ny = y + (d / 3) % 3 - 1, nz = z + d / 9 - 1;
            const int64_t nc = g0.find(nx, ny, nz);
            if (nc >= 0 && idx0[nc] >= 0) m |= 1u << d;
        }
        mask[c] = m;
    }
}

// ---- batched clouds ----
// per-cloud bounding boxes: one block per (cloud, chunk) tile, block
// reduction, one set of atomics per block into that cloud's accumulator
__global__ __launch_bounds__(256) void k_bbox_clouds(PointSrc src, const int32_t* tiles, const int64_t* pt_off,
                                                     BBoxAccum* acc) {
    const int cl = tiles[2 * blockIdx.x];
    const int64_t b0 = pt_off[cl] + (int64_t)tiles[2 * blockIdx.x + 1];
    const int64_t b1 = min(b0 + (int64_t)BBOX_CHUNK, pt_off[cl + 1]);
    float mn0 = __builtin_inff(), mn1 = __builtin_inff(), mn2 = __builtin_inff();
    float mx0 = -__builtin_inff(), mx1 = -__builtin_inff(), mx2 = -__builtin_inff();
    unsigned long long cnt = 0;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        float x, y, z;
        if (!gvdbx_point(src, i, 0.0f, 0.0f, 0.0f, x, y, z)) continue;
        cnt++;
        mn0 = fminf(mn0, x); mn1 = fminf(mn1, y); mn2 = fminf(mn2, z);
        mx0 = fmaxf(mx0, x); mx1 = fmaxf(mx1, y); mx2 = fmaxf(mx2, z);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        mn0 = fminf(mn0, __shfl_xor(mn0, d, 64));
        mn1 = fminf(mn1, __shfl_xor(mn1, d, 64));
        mn2 = fminf(mn2, __shfl_xor(mn2, d, 64));
        mx0 = fmaxf(mx0, __shfl_xor(mx0, d, 64));
        mx1 = fmaxf(mx1, __shfl_xor(mx1, d, 64));
        mx2 = fmaxf(mx2, __shfl_xor(mx2, d, 64));
        cnt += __shfl_xor(cnt, d, 64);
    }
    __shared__ float red[4][6];
    __shared__ unsigned long long redc[4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        red[w][0] = mn0; red[w][1] = mn1; red[w][2] = mn2;
        red[w][3] = mx0; red[w][4] = mx1; red[w][5] = mx2;
        redc[w] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
            mn0 = fminf(mn0, red[i][0]); mn1 = fminf(mn1, red[i][1]); mn2 = fminf(mn2, red[i][2]);
            mx0 = fmaxf(mx0, red[i][3]); mx1 = fmaxf(mx1, red[i][4]); mx2 = fmaxf(mx2, red[i][5]);
            cnt += redc[i];
        }
        if (cnt > 0) {
            BBoxAccum* a = acc + cl;
            atomicMin(&a->mn[0], __float_as_int(mn0));
            atomicMin(&a->mn[1], __float_as_int(mn1));
            atomicMin(&a->mn[2], __float_as_int(mn2));
            atomicMax(&a->mx[0], __float_as_int(mx0));
            atomicMax(&a->mx[1], __float_as_int(mx1));
            atomicMax(&a->mx[2], __float_as_int(mx2));
            atomicAdd(&a->nvalid, cnt);
        }
    }
}

__global__ void k_bbox_clouds_init(BBoxAccum* acc, int ncloud) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncloud; c += gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; a++) {
            acc[c].mn[a] = 0x7F800000;
            acc[c].mx[a] = -1;
        }
        acc[c].nvalid = 0;
    }
}

// cell_cloud over each cloud's cell box, one block per (cloud, chunk) tile
__global__ __launch_bounds__(256) void k_paint_clouds(CellGrid g0, const int32_t* boxes, const int32_t* tiles,
                                                      int32_t* cell_cloud) {
    const int cl = tiles[2 * blockIdx.x];
    const int64_t c0 = tiles[2 * blockIdx.x + 1];
    const int32_t* b = boxes + 6 * cl;
    const int64_t nx = b[3] - b[0], ny = b[4] - b[1], nz = b[5] - b[2];
    const int64_t nb = nx * ny * nz;
    for (int64_t t = c0 + threadIdx.x; t < min(c0 + (int64_t)CELL_CHUNK, nb); t += blockDim.x) {
        const int x = b[0] + (int)(t / (ny * nz)), y = b[1] + (int)((t / nz) % ny), z = b[2] + (int)(t % nz);
        if (const int64_t c = g0.find(x, y, z); c >= 0) cell_cloud[c] = cl;
    }
}

__global__ __launch_bounds__(256) void k_leaf_clouds(const GNode* leaves, int64_t nleaf, CellGrid g0, int logr0,
                                                     const int32_t* cell_cloud, int32_t* leaf_cloud) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nleaf; k += (int64_t)gridDim.x * blockDim.x) {
        const int x = leaves[k].mPos[0] >> logr0, y = leaves[k].mPos[1] >> logr0, z = leaves[k].mPos[2] >> logr0;
        const int64_t c = g0.find(x, y, z);
        leaf_cloud[k] = c >= 0 ? cell_cloud[c] : -1;
    }
}

// per-cloud leaf bounds; leaves come in key order, so a wave's leaves mostly
// share a cloud: the wave reduces when they all do, else lanes go alone
__global__ __launch_bounds__(256) void k_leaf_bounds_clouds(const GNode* leaves, int64_t nleaf, int range0,
                                                            const int32_t* leaf_cloud, int32_t* out6) {
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < nleaf; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = base + threadIdx.x;
        int cl = -1;
        int v[6] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
        if (k < nleaf && leaves[k].mFlags) {
            cl = leaf_cloud[k];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                v[a] = leaves[k].mPos[a];
                v[3 + a] = leaves[k].mPos[a] + range0;
            }
        }
        const int c0 = __shfl(cl, 0, 64);
        const bool same = __ballot(cl == c0 || cl < 0) == ~0ull;
        if (same) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    v[a] = min(v[a], __shfl_xor(v[a], d, 64));
                    v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], d, 64));
                }
            if (lane_id() == 0 && c0 >= 0)
                for (int a = 0; a < 3; a++) {
                    atomicMin(&out6[6 * c0 + a], v[a]);
                    atomicMax(&out6[6 * c0 + 3 + a], v[3 + a]);
                }
        } else if (cl >= 0) {
            for (int a = 0; a < 3; a++) {
                atomicMin(&out6[6 * cl + a], v[a]);
                atomicMax(&out6[6 * cl + 3 + a], v[3 + a]);
            }
        }
    }
}

inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace

int64_t scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void launch_bbox_init(BBoxAccum* acc, hipStream_t s) { hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, s, acc); }

int bbox_parts(int64_t n) { return (int)grid_for(n, 256 * BBOX_UNROLL, BBOX_BLOCKS); }

void launch_bbox(PointSrc src, int64_t n, float tx, float ty, float tz, BBoxAccum* acc, BBoxAccum* parts,
                 hipStream_t s) {
    const int np = bbox_parts(n);
    hipLaunchKernelGGL(k_bbox, dim3(np), dim3(256), 0, s, src, n, tx, ty, tz, parts);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, parts, np, acc);
}

void launch_mark_boxes(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid g0, int range0,
                       float radius, uint8_t* all0, hipStream_t s) {
    hipLaunchKernelGGL(k_mark_boxes, dim3(grid_for(n, 256)), dim3(256), 0, s, src, n, tx, ty, tz, g0, range0, radius,
                       all0);
}

void launch_expand_reach(CellGrid g0, const uint32_t* reach0, uint8_t* base0, uint8_t* all0, int generic,
                         const int32_t* allowed, hipStream_t s) {
    hipLaunchKernelGGL(k_expand_reach, dim3(grid_for(g0.count(), 256)), dim3(256), 0, s, g0, reach0, base0, all0,
                       generic, allowed);
}

void launch_level_up(CellGrid gc, const uint8_t* base_c, const uint8_t* all_c, CellGrid gp, int logres_p,
                     uint8_t* base_p, uint8_t* all_p, hipStream_t s) {
    hipLaunchKernelGGL(k_level_up, dim3(grid_for(gc.count(), 256)), dim3(256), 0, s, gc, base_c, all_c, gp, logres_p,
                       base_p, all_p);
}

void launch_flag_scan_reduce(const uint8_t* base, const uint8_t* all, int64_t n, U2* partials, U2* totals,
                             hipStream_t s) {
    int64_t tiles = scan_tiles(n);
    if (tiles < 1) tiles = 1;
    hipLaunchKernelGGL(k_flag_reduce, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, base, all, n, partials);
    hipLaunchKernelGGL(k_scan_partials<U2>, dim3(1), dim3(1024), 0, s, partials, tiles, totals);
}

void launch_flag_scan_apply(FlagApplyBatch& fb, hipStream_t s) {
    fb.tile_start[0] = 0;
    for (int i = 0; i < fb.n; i++) fb.tile_start[i + 1] = fb.tile_start[i] + std::max<int64_t>(scan_tiles(fb.lv[i].n), 1);
    if (fb.n > 0)
        hipLaunchKernelGGL(k_flag_apply, dim3((unsigned)fb.tile_start[fb.n]), dim3(SCAN_THREADS), 0, s, fb);
}

// Clears every buffer of the batch (byte counts; 4-byte words where aligned).
__global__ __launch_bounds__(256) void k_zero_batch(ZeroBatch zb) {
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, st = (int64_t)gridDim.x * blockDim.x;
    for (int k = 0; k < zb.n; k++) {
        uint8_t* p = zb.p[k];
        const int64_t nb = zb.bytes[k], nw = ((uintptr_t)p & 3) ? 0 : nb >> 2;
        for (int64_t i = t0; i < nw; i += st) ((uint32_t*)p)[i] = 0u;
        for (int64_t i = nw * 4 + t0; i < nb; i += st) p[i] = 0;
    }
}

void launch_zero_batch(const ZeroBatch& zb, hipStream_t s) {
    int64_t mx = 0;
    for (int k = 0; k < zb.n; k++) mx = std::max(mx, zb.bytes[k] / 4 + 1);
    if (zb.n > 0) hipLaunchKernelGGL(k_zero_batch, dim3(grid_for(mx, 256)), dim3(256), 0, s, zb);
}

void launch_count_scan_reduce(const int32_t* cnt, int64_t n, uint32_t* partials, uint32_t* total, hipStream_t s) {
    int64_t tiles = scan_tiles(n + 1);
    hipLaunchKernelGGL(k_count_reduce, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, cnt, n, partials);
    hipLaunchKernelGGL(k_scan_partials<uint32_t>, dim3(1), dim3(1024), 0, s, partials, tiles, total);
}

void launch_count_scan_apply(const int32_t* cnt, int64_t n, const uint32_t* partials, int32_t* off,
                             hipStream_t s) {
    int64_t tiles = scan_tiles(n + 1);
    if (((uintptr_t)cnt | (uintptr_t)off) % 16 == 0)
        hipLaunchKernelGGL(k_count_apply<true>, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, cnt, n, partials, off);
    else
        hipLaunchKernelGGL(k_count_apply<false>, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, cnt, n, partials, off);
}

void launch_checksum(const uint32_t* p, int64_t n, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, s, out, (int64_t)2, 0ull);
    if (n > 0) hipLaunchKernelGGL(k_checksum, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, out);
}

void launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_f32(float* p, int64_t n, float v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}

void launch_link(const LinkBatch& lb, hipStream_t s) {
    if (lb.n > 0 && lb.start[lb.n] > 0)
        hipLaunchKernelGGL(k_link, dim3(grid_for(lb.start[lb.n], 256)), dim3(256), 0, s, lb);
}

void launch_clist_init(const ClistInit& ci, hipStream_t s) {
    hipLaunchKernelGGL(k_clist_init, dim3(grid_for(std::max<int64_t>(ci.start[ci.n], 1), 256)), dim3(256), 0, s, ci);
}

void launch_bounds_init(int32_t* out6, hipStream_t s) { hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, out6); }

void launch_atlas_assign(GNode* leaves, int64_t nleaf, int sx, int sy, int brick_res, int apron, GAtlasNode* amap,
                         int64_t nmap, hipStream_t s) {
    if (nmap > 0)
        hipLaunchKernelGGL(k_atlas_assign, dim3(grid_for(nmap, 256)), dim3(256), 0, s, leaves, nleaf, sx, sy, brick_res,
                           apron, amap, nmap);
}

void launch_leaf_bounds(const GNode* leaves, int64_t nleaf, int range0, int32_t* out6, hipStream_t s) {
    hipLaunchKernelGGL(k_leaf_bounds, dim3(grid_for(nleaf, 256, 256)), dim3(256), 0, s, leaves, nleaf, range0, out6);
}

void launch_reach_sorted(const float4* sorted, const int32_t* total, int64_t nmax, ReachParams rp, uint32_t* reach0,
                         hipStream_t s) {
    if (nmax > 0)
        hipLaunchKernelGGL(k_reach_sorted, dim3(grid_for(nmax, 256 * REACH_UNROLL)), dim3(256), 0, s, sorted, total, nmax, rp,
                           reach0);
}

void launch_subcell_len(const float4* sorted, const int32_t* total, int64_t nmax, BinParams bp,
                        unsigned long long* len, hipStream_t s) {
    if (nmax > 0)
        hipLaunchKernelGGL(k_subcell_len, dim3(grid_for(nmax, 256, 2048)), dim3(256), 0, s, sorted, total, nmax, bp, len);
}

void launch_bin_keys(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid gb, int logr0, CloudMap cm,
                     uint32_t* key, uint32_t* val, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_bin_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, src, n, tx, ty, tz, gb, logr0, cm, key,
                           val);
}

void launch_bin_gather(PointSrc src, int64_t n, float tx, float ty, float tz, const uint32_t* key,
                       const uint32_t* val, uint32_t ncell, float4* sorted, const ReachParams* rp, uint32_t* reach0,
                       hipStream_t s) {
    const ReachParams r = rp ? *rp : ReachParams{};
    if (n > 0)
        hipLaunchKernelGGL(k_bin_gather, dim3(grid_for(n, 256 * BIN_UNROLL)), dim3(256), 0, s, src, n, tx, ty, tz, key,
                           val, ncell, sorted, r, rp ? reach0 : nullptr);
}

void launch_bin_offsets(const uint32_t* key, int64_t n, int64_t ncell, int32_t* off, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_offsets, dim3(grid_for(ncell + 1, 256)), dim3(256), 0, s, key, n, ncell, off);
}

void launch_leaf_nbmask(const int32_t* idx0, CellGrid g0, uint32_t* mask, unsigned long long* reset, hipStream_t s) {
    hipLaunchKernelGGL(k_leaf_nbmask, dim3(grid_for(std::max<int64_t>(g0.count(), 1), 256)), dim3(256), 0, s, idx0, g0,
                       mask, reset);
}


void launch_bbox_clouds(PointSrc src, const int32_t* tiles, int64_t ntiles, const int64_t* pt_off, BBoxAccum* acc,
                        hipStream_t s) {
    if (ntiles > 0) hipLaunchKernelGGL(k_bbox_clouds, dim3((unsigned)ntiles), dim3(256), 0, s, src, tiles, pt_off, acc);
}

void launch_bbox_clouds_init(BBoxAccum* acc, int ncloud, hipStream_t s) {
    if (ncloud > 0) hipLaunchKernelGGL(k_bbox_clouds_init, dim3(grid_for(ncloud, 256)), dim3(256), 0, s, acc, ncloud);
}

void launch_paint_clouds(CellGrid g0, const int32_t* boxes, const int32_t* tiles, int64_t ntiles, int32_t* cell_cloud,
                         hipStream_t s) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_paint_clouds, dim3((unsigned)ntiles), dim3(256), 0, s, g0, boxes, tiles, cell_cloud);
}

void launch_leaf_clouds(const GNode* leaves, int64_t nleaf, CellGrid g0, int logr0, const int32_t* cell_cloud,
                        int32_t* leaf_cloud, hipStream_t s) {
    if (nleaf > 0)
        hipLaunchKernelGGL(k_leaf_clouds, dim3(grid_for(nleaf, 256)), dim3(256), 0, s, leaves, nleaf, g0, logr0,
                           cell_cloud, leaf_cloud);
}

void launch_leaf_bounds_clouds(const GNode* leaves, int64_t nleaf, int range0, const int32_t* leaf_cloud,
                               int32_t* out6, hipStream_t s) {
    if (nleaf > 0)
        hipLaunchKernelGGL(k_leaf_bounds_clouds, dim3(grid_for(nleaf, 256)), dim3(256), 0, s, leaves, nleaf, range0,
                           leaf_cloud, out6);
}
