// incre.hip -- gfx950 kernels of the incremental topology update
// (ActivateIncreBricksGPU, gvdb_volume_gvdb.cpp:1020-1155; AccumulateTopology
// 1158-1178).
//
// The reference finds a box cell's node by descending the child lists from
// the root (getPosParent, cuda_gvdb_nodes.cuh:303-348), collects the missing
// cells as 64-bit keys, sorts them, appends the nodes on the host, links
// them by descending again (gvdbLinkBricks) and delinks the leaves no point
// reached plus the inner nodes left childless (gvdbDelinkLeafBricks /
// gvdbDelinkBricks, cuda_gvdb_nodes.cuh:401-467).  Here every level keeps a
// dense map cell -> reachable node over a grid that holds every cell the
// update can touch, so each lookup is one load:
//   1. k_incre_mark: per point and level, the box cells (clipped to the
//      root); a cell with a reachable node marks it (leaves), a cell without
//      one sets its flag in newf;
//   2. the ordered flag scan appends the missing nodes in key order (the
//      topology code's k_flag_apply with a pool offset) and enters them in
//      the map, so the map is what a descent would find after the appends;
//   3. the link kernel of the topology code, restricted to flagged nodes;
//   4. k_incre_delink per level, leaves first: parents come from the maps of
//      the level above, which no delink has touched yet, exactly as the
//      reference's descent at that point;
//   5. k_incre_final_idx drops the delinked nodes from the maps.
#include "gvdbx_kernels.h"

namespace {

inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

__global__ __launch_bounds__(256) void k_remap_idx(CellGrid from, const int32_t* src, CellGrid to, int32_t* dst) {
    const int64_t n = to.count();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        to.cell(i, x, y, z);
        const int64_t f = src ? from.find(x, y, z) : -1;
        dst[i] = f >= 0 ? src[f] : -1;
    }
}

// gvdbCalcIncreExtraBrickId (cuda_gvdb_particles.cuh:420-465): the box
// [int(w) - r, int(w) + r], corners divided by the level's range with C
// truncation, walked at every level below the root.
__global__ __launch_bounds__(256) void k_incre_mark(PointSrc src, int64_t n, float tx, float ty, float tz,
                                                    float radius, IncreMark im) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float w[3];
        if (!gvdbx_point(src, i, tx, ty, tz, w[0], w[1], w[2])) continue;
        int bmin[3], bmax[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float f = (float)(int)w[a];
            bmin[a] = (int)(f - radius);
            bmax[a] = (int)(f + radius);
        }
        for (int lev = 0; lev < im.depth; lev++) {
            const int rres = (int)im.range[lev];
            int lo[3], hi[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                // cells outside the root are not emitted
                lo[a] = max(bmin[a] / rres, im.root_cell[lev][a]);
                hi[a] = min(bmax[a] / rres, im.root_cell[lev][a] + im.root_span[lev] - 1);
            }
            const CellGrid& g = im.g[lev];
            for (int x = lo[0]; x <= hi[0]; x++)
                for (int y = lo[1]; y <= hi[1]; y++)
                    for (int z = lo[2]; z <= hi[2]; z++) {
                        const int64_t c = g.find(x, y, z);
                        if (c < 0) continue;   // the grid holds every touched cell
                        const int32_t id = im.idx[lev][c];
                        if (id < 0) im.newf[lev][c] = 1;
                        else if (lev == 0 && id < im.nleaf) im.marker0[id] = 1;
                    }
        }
    }
}

__global__ __launch_bounds__(256) void k_incre_leaf_flags(GNode* leaves, int64_t n, const uint8_t* marker, int accum) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t f = (uint8_t)((accum ? leaves[k].mFlags : 0) | marker[k]);
        leaves[k].mFlags = f;
    }
}

// gvdbDelinkLeafBricks (lev 0) / gvdbDelinkBricks (lev >= 1)
__global__ __launch_bounds__(256) void k_incre_delink(DelinkParams dp) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < dp.n; k += (int64_t)gridDim.x * blockDim.x) {
        GNode& nd = dp.nodes[k];
        if (dp.lev == 0) {
            if (nd.mFlags || nd.mParent == GVDBX_ID_UNDEF64) continue;
        } else {
            if (nd.mParent == GVDBX_ID_UNDEF64) {
                nd.mFlags = 0;
                continue;
            }
            const uint64_t* own = dp.own_clist + k * dp.own_len;
            bool any = false;
            for (int64_t c = 0; c < dp.own_len && !any; c++) any = own[c] != GVDBX_ID_UNDEF64;
            if (any) {
                nd.mFlags = 1;
                continue;
            }
            nd.mFlags = 0;
        }
        int64_t p = 0;
        if (!dp.parent_is_root) {
            int pc[3];
#pragma unroll
            for (int a = 0; a < 3; a++) pc[a] = (int)(nd.mPos[a] / dp.range_p);
            const int64_t pcell = dp.gp.find(pc[0], pc[1], pc[2]);
            if (pcell < 0) continue;
            p = dp.idx_p[pcell];
        }
        if (p < 0 || p >= dp.nparent) continue;
        const GNode& pn = dp.parent_nodes[p];
        int64_t pin[3];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            pin[a] = ((int64_t)(nd.mPos[a] - pn.mPos[a]) * dp.res_p) / dp.range_p;
            ok = ok && pin[a] >= 0 && pin[a] < dp.res_p;
        }
        nd.mParent = GVDBX_ID_UNDEF64;
        if (!ok) continue;
        const int64_t len = (int64_t)dp.res_p * dp.res_p * dp.res_p;
        const int64_t flat = (int64_t)(pn.mChildList >> 16) * len + (pin[2] * dp.res_p + pin[1]) * dp.res_p + pin[0];
        if (flat >= 0 && flat < dp.nparent * len) dp.parent_clist[flat] = GVDBX_ID_UNDEF64;
    }
}

__global__ __launch_bounds__(256) void k_incre_final_idx(int64_t n, int32_t* idx, const GNode* nodes) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = idx[i];
        if (k >= 0 && nodes[k].mParent == GVDBX_ID_UNDEF64) idx[i] = -1;
    }
}

__global__ __launch_bounds__(256) void k_leaf_flags32(const GNode* leaves, int64_t n, int32_t* flags) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        flags[k] = leaves[k].mFlags ? 1 : 0;
}

// UpdateAtlas (VG.cpp:2630-2704) when some leaves are inactive: AtlasAlloc
// hands out slots in leaf order to the active leaves only.
__global__ __launch_bounds__(256) void k_atlas_clear_map(GAtlasNode* amap, int64_t nmap) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nmap; k += (int64_t)gridDim.x * blockDim.x) {
        GAtlasNode an;
        an.mPos[0] = an.mPos[1] = an.mPos[2] = (int32_t)GVDBX_ID_UNDEFL;
        an.mLeafID = (int32_t)GVDBX_ID_UNDEFL;
        amap[k] = an;
    }
}

__global__ __launch_bounds__(256) void k_atlas_assign_rank(GNode* leaves, int64_t nleaf, const int32_t* rank, int sx,
                                                           int sy, int br, int apron, GAtlasNode* amap, int64_t nmap) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nleaf; k += (int64_t)gridDim.x * blockDim.x) {
        GNode& nd = leaves[k];
        if (!nd.mFlags) continue;
        const int64_t slot = rank[k];
        if (slot >= nmap) continue;
        const int64_t a2 = (int64_t)sx * sy;
        const int64_t z = slot / a2, rem = slot - z * a2;
        const int64_t y = rem / sx, x = rem - y * sx;
        nd.mValue[0] = (int32_t)(x * br + apron);
        nd.mValue[1] = (int32_t)(y * br + apron);
        nd.mValue[2] = (int32_t)(z * br + apron);
        GAtlasNode an;
        an.mPos[0] = nd.mPos[0]; an.mPos[1] = nd.mPos[1]; an.mPos[2] = nd.mPos[2];
        an.mLeafID = (int32_t)k;
        amap[slot] = an;
    }
}

}  // namespace

void launch_remap_idx(CellGrid from, const int32_t* src, CellGrid to, int32_t* dst, hipStream_t s) {
    const int64_t n = to.count();
    if (n > 0) hipLaunchKernelGGL(k_remap_idx, dim3(grid_for(n, 256)), dim3(256), 0, s, from, src, to, dst);
}

void launch_incre_mark(PointSrc src, int64_t n, float tx, float ty, float tz, float radius, const IncreMark& im,
                       hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_incre_mark, dim3(grid_for(n, 256)), dim3(256), 0, s, src, n, tx, ty, tz, radius, im);
}

void launch_incre_leaf_flags(GNode* leaves, int64_t n, const uint8_t* marker, int accum, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_incre_leaf_flags, dim3(grid_for(n, 256)), dim3(256), 0, s, leaves, n, marker, accum);
}

void launch_incre_delink(const DelinkParams& dp, hipStream_t s) {
    if (dp.n > 0) hipLaunchKernelGGL(k_incre_delink, dim3(grid_for(dp.n, 256)), dim3(256), 0, s, dp);
}

void launch_incre_final_idx(CellGrid g, int32_t* idx, const GNode* nodes, hipStream_t s) {
    const int64_t n = g.count();
    if (n > 0) hipLaunchKernelGGL(k_incre_final_idx, dim3(grid_for(n, 256)), dim3(256), 0, s, n, idx, nodes);
}

void launch_leaf_flags32(const GNode* leaves, int64_t n, int32_t* flags, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_leaf_flags32, dim3(grid_for(n, 256)), dim3(256), 0, s, leaves, n, flags);
}

void launch_atlas_assign_rank(GNode* leaves, int64_t nleaf, const int32_t* rank, int sx, int sy, int brick_res,
                              int apron, GAtlasNode* amap, int64_t nmap, hipStream_t s) {
    if (nmap > 0) hipLaunchKernelGGL(k_atlas_clear_map, dim3(grid_for(nmap, 256)), dim3(256), 0, s, amap, nmap);
    if (nleaf > 0)
        hipLaunchKernelGGL(k_atlas_assign_rank, dim3(grid_for(nleaf, 256)), dim3(256), 0, s, leaves, nleaf, rank, sx, sy,
                           brick_res, apron, amap, nmap);
}
