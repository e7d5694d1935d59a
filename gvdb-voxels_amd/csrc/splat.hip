// splat.hip -- gfx950 kernels that turn binned points into voxel values.
//
// Every kernel works on one 8x8x8 tile of a leaf brick per wavefront: lane
// (x,y) owns the z-column of the tile and keeps its eight voxels in
// registers.  Candidate points come from the 27 neighbouring bins of the
// brick (points sorted by home brick), are filtered to those whose support
// reaches the tile, and are staged in LDS.  Results are written once to the
// brick-linear atlas (leaf k at k*brick_res^3 floats).  No global float
// atomics are used; the cubic splat's LDS adds come from the tile's single
// wave in an order fixed by the data, so every result is deterministic.
#include <algorithm>
#include <climits>
#include <type_traits>
#include <cmath>

#include "gvdbx_kernels.h"

namespace {

constexpr int CAND = 512;   // staged candidates per wavefront

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float rflf(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

__device__ __forceinline__ int tile_dim(int res0) { return res0 < 8 ? res0 : 8; }

// Position of tile t inside its brick.
__device__ __forceinline__ void tile_origin(int res0, int64_t t, int& ox, int& oy, int& oz) {
    int td = tile_dim(res0);
    int per = res0 / td;
    ox = (int)(t % per) * td;
    oy = (int)((t / per) % per) * td;
    oz = (int)(t / ((int64_t)per * per)) * td;
}

// ---------------------------------------------------------------------------
// GatherDensity / GatherLevelSet (cuda_gvdb_particles.cuh:696-773) in scatter
// form.  Semantics: a point belongs to the list of subcell s when s lies in
// the point's subcell box (gvdbInsertSubcell, cuda_gvdb_particles.cuh:241-270);
// a voxel takes min (level set) or max (density) over the members of its
// subcell.  Because sqrt and the subtraction of r are monotone, the min of
// |p-v|-r equals sqrt(min |p-v|^2) - r bit for bit, so the kernel keeps one
// min of squared distance per voxel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_gather(GatherParams gp, int res0, int tiles_per_brick, int64_t nleaf_tiles) {
    __shared__ float4 cand[CAND];
    const int lane = threadIdx.x;
    const int64_t tile = gvdbx_xcd_item(blockIdx.x, nleaf_tiles, gp.xcd_remap);
    if (tile < 0) return;
    const int64_t k = tile / tiles_per_brick;
    if (k >= gp.nleaf) return;
    const int td = tile_dim(res0);
    int tox, toy, toz;
    tile_origin(res0, tile % tiles_per_brick, tox, toy, toz);
    const GNode nd = gp.leaves[k];
    int shx = 0, shy = 0, shz = 0;   // cloud-local voxel coordinates (batched clouds)
    if (gp.leaf_cloud) {
        const int cl = gp.leaf_cloud[k];
        shx = gp.cell_off[3 * cl] * res0, shy = gp.cell_off[3 * cl + 1] * res0, shz = gp.cell_off[3 * cl + 2] * res0;
    }
    const int T0x = nd.mPos[0] - shx + tox, T0y = nd.mPos[1] - shy + toy, T0z = nd.mPos[2] - shz + toz;
    const int lx = lane % td, ly = lane / td;
    const bool lane_ok = ly < td;
    const float vx = (float)(T0x + lx), vy = (float)(T0y + ly);
    float vz[8];
    float m[8];
#pragma unroll
    for (int z = 0; z < 8; z++) {
        vz[z] = (float)(T0z + z);
        m[z] = __builtin_inff();
    }
    const int cx = nd.mPos[0] / res0, cy = nd.mPos[1] / res0, cz = nd.mPos[2] / res0;
    const float R = gp.box_radius;
    const int sc = gp.sc, scs = gp.sc_shift;
    const float cut = gp.cutoff;
    int cnt = 0;

    auto process = [&](int count) {
        __syncthreads();
        for (int j = 0; j < count; j++) {
            float4 c = cand[j];
            uint32_t pk = rfl(__float_as_uint(c.w));
            float px = rflf(c.x), py = rflf(c.y), pz = rflf(c.z);
            int xlo = pk & 7, xhi = (pk >> 3) & 7, ylo = (pk >> 6) & 7, yhi = (pk >> 9) & 7;
            int zlo = (pk >> 12) & 7, zhi = (pk >> 15) & 7;
            if (lane_ok && lx >= xlo && lx <= xhi && ly >= ylo && ly <= yhi) {
                float dx = vx - px, dy = vy - py;
                float dxy = __builtin_fmaf(dy, dy, dx * dx);
#pragma unroll
                for (int z = 0; z < 8; z++) {
                    if (z >= zlo && z <= zhi) {
                        float dz = vz[z] - pz;
                        m[z] = fminf(m[z], __builtin_fmaf(dz, dz, dxy));
                    }
                }
            }
        }
        __syncthreads();
    };

    for (int dcx = -1; dcx <= 1; dcx++)
        for (int dcy = -1; dcy <= 1; dcy++)
            for (int dcz = -1; dcz <= 1; dcz++) {
                int bx = cx + dcx, by = cy + dcy, bz = cz + dcz;
                const int64_t ci = gp.gb.find(bx, by, bz);
                if (ci < 0) continue;
                int beg = gp.off[ci], end = gp.off[ci + 1];
                for (int j0 = beg; j0 < end; j0 += 64) {
                    int j = j0 + lane;
                    bool ok = j < end;
                    float4 p = make_float4(0, 0, 0, 0);
                    uint32_t pk = 0;
                    if (ok) {
                        p = gp.sorted[j];
                        float pw[3] = {p.x, p.y, p.z};
                        int T0[3] = {T0x, T0y, T0z};
                        int lo_r[3], hi_r[3];
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            // subcell box (cuda_gvdb_particles.cuh:237-242) as a voxel range
                            int lo = gvdbx_cdiv_mul((int)(pw[a] - R), scs);
                            int hi = gvdbx_cdiv_mul((int)(pw[a] + 1 + R), scs) + sc - 1;
                            lo = max(lo, (int)ceilf(pw[a] - cut));
                            hi = min(hi, (int)floorf(pw[a] + cut));
                            lo_r[a] = max(lo - T0[a], 0);
                            hi_r[a] = min(hi - T0[a], td - 1);
                            ok = ok && lo_r[a] <= hi_r[a];
                        }
                        pk = (uint32_t)lo_r[0] | ((uint32_t)hi_r[0] << 3) | ((uint32_t)lo_r[1] << 6) |
                             ((uint32_t)hi_r[1] << 9) | ((uint32_t)lo_r[2] << 12) | ((uint32_t)hi_r[2] << 15);
                    }
                    uint64_t bal = __ballot(ok);
                    if (ok) {
                        int pos = cnt + (int)__popcll(bal & ((1ull << lane) - 1ull));
                        cand[pos] = make_float4(p.x, p.y, p.z, __uint_as_float(pk));
                    }
                    cnt += (int)__popcll(bal);
                    if (cnt > CAND - 64) {
                        process(cnt);
                        cnt = 0;
                    }
                }
            }
    if (cnt > 0) process(cnt);

    if (!lane_ok) return;
    const int br = gp.brick_res, a = gp.apron;
    float* brick = gp.chan + (int64_t)k * br * br * br;
    const float r = gp.radius;
#pragma unroll
    for (int z = 0; z < 8; z++) {
        if (z < td) {
            int64_t vi = ((int64_t)(toz + z + a) * br + (toy + ly + a)) * br + (tox + lx + a);
            float init = gp.accum ? brick[vi] : (gp.mode == 1 ? 3.0f : 0.0f);
            float c = sqrtf(m[z]);
            float val = gp.mode == 1 ? fminf(init, c - r) : fmaxf(init, r - c);
            brick[vi] = val;
        }
    }
}

// ---------------------------------------------------------------------------
// Same semantics as k_gather, point-centric: the staged members are spread
// over the lanes, and each lane walks its member's voxel box (subcell box,
// cutoff box and tile intersected) row by row, skipping the rows and the x
// spans outside the cutoff sphere, and folds |p-v|^2 into the tile's LDS copy
// with an integer min (non-negative floats order as their bit patterns).  The
// min is exact and order-free, so the result is bit-identical to k_gather.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_gather_pc(GatherParams gp, int res0, int tiles_per_brick, int64_t nleaf_tiles) {
    __shared__ float4 cand[CAND];
    __shared__ uint32_t acc[512];
    const int lane = threadIdx.x;
    const int64_t tile = gvdbx_xcd_item(blockIdx.x, nleaf_tiles, gp.xcd_remap);
    if (tile < 0) return;
    const int64_t k = tile / tiles_per_brick;
    if (k >= gp.nleaf) return;
    const int td = tile_dim(res0);
    const int nv = td * td * td;
    int tox, toy, toz;
    tile_origin(res0, tile % tiles_per_brick, tox, toy, toz);
    const GNode nd = gp.leaves[k];
    int shx = 0, shy = 0, shz = 0;   // cloud-local voxel coordinates (batched clouds)
    if (gp.leaf_cloud) {
        const int cl = gp.leaf_cloud[k];
        shx = gp.cell_off[3 * cl] * res0, shy = gp.cell_off[3 * cl + 1] * res0, shz = gp.cell_off[3 * cl + 2] * res0;
    }
    const int T0x = nd.mPos[0] - shx + tox, T0y = nd.mPos[1] - shy + toy, T0z = nd.mPos[2] - shz + toz;
    for (int i = lane; i < nv; i += 64) acc[i] = 0x7F800000u;   // +inf
    const int cx = nd.mPos[0] / res0, cy = nd.mPos[1] / res0, cz = nd.mPos[2] / res0;
    const float R = gp.box_radius;
    const int sc = gp.sc, scs = gp.sc_shift;
    const float cut = gp.cutoff;
    // row pruning radius: a little beyond the cutoff (the per-voxel values are
    // exact; pruning only has to keep every voxel the cutoff keeps)
    const float cutm = (cut + 0.01f) * (cut + 0.01f);
    int cnt = 0;

    auto process = [&](int count) {
        __syncthreads();
        for (int j = lane; j < count; j += 64) {
            const float4 c = cand[j];
            const uint32_t pk = __float_as_uint(c.w);
            const float px = c.x, py = c.y, pz = c.z;
            const int xlo = pk & 7, xhi = (pk >> 3) & 7, ylo = (pk >> 6) & 7, yhi = (pk >> 9) & 7;
            const int zlo = (pk >> 12) & 7, zhi = (pk >> 15) & 7;
            const float xc = px - (float)T0x;
            for (int z = zlo; z <= zhi; z++) {
                const float dz = (float)(T0z + z) - pz;
                const float dz2 = dz * dz;
                if (dz2 > cutm) continue;
                for (int y = ylo; y <= yhi; y++) {
                    const float dy = (float)(T0y + y) - py;
                    const float row = dz2 + dy * dy;
                    if (row > cutm) continue;
                    const float half = sqrtf(cutm - row);
                    const int xa = max(xlo, (int)ceilf(xc - half)), xb = min(xhi, (int)floorf(xc + half));
                    uint32_t* arow = acc + (z * td + y) * td;
                    for (int x = xa; x <= xb; x++) {
                        const float dx = (float)(T0x + x) - px;
                        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        atomicMin(&arow[x], __float_as_uint(d2));
                    }
                }
            }
        }
        __syncthreads();
    };

    for (int dcx = -1; dcx <= 1; dcx++)
        for (int dcy = -1; dcy <= 1; dcy++)
            for (int dcz = -1; dcz <= 1; dcz++) {
                int bx = cx + dcx, by = cy + dcy, bz = cz + dcz;
                const int64_t ci = gp.gb.find(bx, by, bz);
                if (ci < 0) continue;
                int beg = gp.off[ci], end = gp.off[ci + 1];
                for (int j0 = beg; j0 < end; j0 += 64) {
                    int j = j0 + lane;
                    bool ok = j < end;
                    float4 p = make_float4(0, 0, 0, 0);
                    uint32_t pk = 0;
                    if (ok) {
                        p = gp.sorted[j];
                        float pw[3] = {p.x, p.y, p.z};
                        int T0[3] = {T0x, T0y, T0z};
                        int lo_r[3], hi_r[3];
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            // subcell box (cuda_gvdb_particles.cuh:237-242) as a voxel range
                            int lo = gvdbx_cdiv_mul((int)(pw[a] - R), scs);
                            int hi = gvdbx_cdiv_mul((int)(pw[a] + 1 + R), scs) + sc - 1;
                            lo = max(lo, (int)ceilf(pw[a] - cut));
                            hi = min(hi, (int)floorf(pw[a] + cut));
                            lo_r[a] = max(lo - T0[a], 0);
                            hi_r[a] = min(hi - T0[a], td - 1);
                            ok = ok && lo_r[a] <= hi_r[a];
                        }
                        pk = (uint32_t)lo_r[0] | ((uint32_t)hi_r[0] << 3) | ((uint32_t)lo_r[1] << 6) |
                             ((uint32_t)hi_r[1] << 9) | ((uint32_t)lo_r[2] << 12) | ((uint32_t)hi_r[2] << 15);
                    }
                    uint64_t bal = __ballot(ok);
                    if (ok) {
                        int pos = cnt + (int)__popcll(bal & ((1ull << lane) - 1ull));
                        cand[pos] = make_float4(p.x, p.y, p.z, __uint_as_float(pk));
                    }
                    cnt += (int)__popcll(bal);
                    if (cnt > CAND - 64) {
                        process(cnt);
                        cnt = 0;
                    }
                }
            }
    if (cnt > 0) process(cnt);
    __syncthreads();
    const int br = gp.brick_res, a = gp.apron;
    float* brick = gp.chan + (int64_t)k * br * br * br;
    const float r = gp.radius;
    for (int i = lane; i < nv; i += 64) {
        const int x = i % td, y = (i / td) % td, z = i / (td * td);
        const int64_t vi = ((int64_t)(toz + z + a) * br + (toy + y + a)) * br + (tox + x + a);
        const float init = gp.accum ? brick[vi] : (gp.mode == 1 ? 3.0f : 0.0f);
        const float c = sqrtf(__uint_as_float(acc[i]));
        brick[vi] = gp.mode == 1 ? fminf(init, c - r) : fmaxf(init, r - c);
    }
}

// ---------------------------------------------------------------------------
// Same semantics as k_gather without LDS: each 64-point chunk of a
// neighbouring bin is filtered lane-parallel, then the surviving candidates
// are walked one at a time with ballot + readlane (their coordinates land in
// scalar registers), and the next chunk's load is in flight meanwhile.
//
// A tile's candidate sequence is the chunks (64 points, the last one of a bin
// possibly shorter) of its 27 neighbour bins, bins in NB_ORDER.  Dense tiles
// (more than 2 split chunks, split a tuning knob) are cut into ns =
// ceil(chunks / split) work items; item s walks chunks s, s + ns, s + 2 ns, ... so that
// every item meets the near bins first.  Items of one tile fold their minima
// into a scratch tile with integer atomicMin (exact and order-free: the
// minima are bit patterns of non-negative floats), and a last pass writes
// the bricks from it with the same arithmetic as a single wave.
// ---------------------------------------------------------------------------
// Subcell membership range of a point along one axis clipped to the cutoff
// (cuda_gvdb_particles.cuh:237-242), the arithmetic of the gather's filter.
__device__ __forceinline__ void gather_range(float p, float R, int sc, int scs, float cut, int& lo_c, int& hi_c) {
    const int lo = gvdbx_cdiv_mul((int)(p - R), scs);
    const int hi = gvdbx_cdiv_mul((int)(p + 1 + R), scs) + sc - 1;
    lo_c = max(lo, (int)ceilf(p - cut));
    hi_c = min(hi, (int)floorf(p + cut));
}

// Each point's ranges once per gather instead of once per tile that reads
// it (up to 27).  Offsets: floor(p) - lo_c + 1 in [0, ceil(R) + sc + 2] and
// hi_c - floor(p) in [0, R + sc + 1] when the range is not empty.
// pm: the position membership is computed from (the point inserted into
// the lists); pc: the position distances are measured from (the same, or
// its FP16 round trip), whose floor the offsets are taken from.
__device__ __forceinline__ uint32_t gather_box_word2(const float (&pm)[3], const float (&pc)[3], float R, int sc,
                                                     int scs, float cut, int (&lo_c)[3], int (&hi_c)[3]) {
    uint32_t bw = 0;
    bool empty = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // membership alone (gather_range's box), then the cutoff from pc
        const int lo = gvdbx_cdiv_mul((int)(pm[a] - R), scs);
        const int hi = gvdbx_cdiv_mul((int)(pm[a] + 1 + R), scs) + sc - 1;
        lo_c[a] = max(lo, (int)ceilf(pc[a] - cut));
        hi_c[a] = min(hi, (int)floorf(pc[a] + cut));
        const int f = (int)floorf(pc[a]);
        empty = empty || lo_c[a] > hi_c[a];
        bw |= ((uint32_t)(f - lo_c[a] + 1) & 31u) << (10 * a);
        bw |= ((uint32_t)(hi_c[a] - f) & 31u) << (10 * a + 5);
    }
    return empty ? 0x80000000u : bw;
}

__device__ __forceinline__ uint32_t gather_box_word(const float4& q, float R, int sc, int scs, float cut, int (&lo_c)[3],
                                                    int (&hi_c)[3]) {
    const float pw[3] = {q.x, q.y, q.z};
    uint32_t bw = 0;
    bool empty = false;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        gather_range(pw[a], R, sc, scs, cut, lo_c[a], hi_c[a]);
        const int f = (int)floorf(pw[a]);
        empty = empty || lo_c[a] > hi_c[a];
        bw |= ((uint32_t)(f - lo_c[a] + 1) & 31u) << (10 * a);
        bw |= ((uint32_t)(hi_c[a] - f) & 31u) << (10 * a + 5);
    }
    return empty ? 0x80000000u : bw;
}

__global__ __launch_bounds__(256) void k_gather_box(const float4* sorted, const int32_t* total_ptr, int64_t nmax,
                                                    float R, int sc, int scs, float cut, uint32_t* box) {
    const int64_t total = min((int64_t)*total_ptr, nmax);
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
        int lo_c[3], hi_c[3];
        box[j] = gather_box_word(sorted[j], R, sc, scs, cut, lo_c, hi_c);
    }
}

// Reach classes.  A point's class has bit a set when its range along axis a
// (membership clipped to the cutoff, as above) reaches past the upper face
// of its bin (level-0 cell).  A tile in cell c needs, from the bin at c + o,
// only the classes that can reach c: bit a set where o_a = -1, bit a clear
// where o_a = +1 (unless a point of that bin reaches past both faces along
// a), anything where o_a = 0.  Each bin's points are grouped by class into
// csorted as (x, y, z, packed ranges), the class starts go to coff[8 ci + k]
// (coff[8 nbin] = the total) and the axes with points reaching past both
// faces to both[ci].  The order inside a bin does not change a gather: its
// result is a minimum (bins of at most 256 entries are grouped with LDS
// atomics, in no fixed order inside a class; larger ones stably).
constexpr int CLS_UNROLL = 4;   // chunks of 64 entries loaded at once

// Entries j0 + 64 u + lane (u < CLS_UNROLL) of a bin ending at `end`: the
// point with its packed ranges in .w and its class (-1: no entry); bb
// collects the axes along which one of them reaches past both faces.  A
// point's bin is the level-0 cell holding it (cells of side 2^logr0), the
// same floor the binning takes, so its faces come from the point itself.
__device__ __forceinline__ void class_load(const float4* sorted, int j0, int end, int logr0, float R, int sc, int scs,
                                           float cut, const Quant16& qz, float4 (&q)[CLS_UNROLL],
                                           int (&cls)[CLS_UNROLL], uint32_t& bb) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int u = 0; u < CLS_UNROLL; u++) {
        const int j = j0 + 64 * u + lane;
        cls[u] = -1;
        if (j < end) {
            q[u] = sorted[j];
            cls[u] = 0;
        }
    }
#pragma unroll
    for (int u = 0; u < CLS_UNROLL; u++) {
        if (cls[u] < 0) continue;
        int lo_c[3], hi_c[3];
        const float pw[3] = {q[u].x, q[u].y, q[u].z};
        uint32_t bw;
        if (qz.on) {
            // FP16 lists: membership from the inserted point, distances from its round trip
            const float pc[3] = {quant16_round_trip(pw[0], qz.mn[0], qz.rg[0]),
                                 quant16_round_trip(pw[1], qz.mn[1], qz.rg[1]),
                                 quant16_round_trip(pw[2], qz.mn[2], qz.rg[2])};
            bw = gather_box_word2(pw, pc, R, sc, scs, cut, lo_c, hi_c);
            q[u].x = pc[0], q[u].y = pc[1], q[u].z = pc[2];
        } else {
            bw = gather_box_word(q[u], R, sc, scs, cut, lo_c, hi_c);
        }
        q[u].w = __uint_as_float(bw);
        if (!(bw >> 31)) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int O = gvdbx_cell(pw[a], logr0) << logr0, res0 = 1 << logr0;
                const bool up = hi_c[a] >= O + res0;
                cls[u] |= (int)up << a;
                bb |= (uint32_t)(up && lo_c[a] < O) << a;
            }
        }
    }
}

// Places the loaded entries of the first nu chunks (the others hold none):
// base[k] is where the next entry of class k goes (advanced past them);
// entries go in order u, then lane.
__device__ __forceinline__ void class_place(const float4 (&q)[CLS_UNROLL], const int (&cls)[CLS_UNROLL], int nu,
                                            int (&base)[8], float4* csorted) {
#pragma unroll
    for (int u = 0; u < CLS_UNROLL; u++) {
        if (u >= nu) break;
        int pos = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint64_t bal = __ballot(cls[u] == k);
            if (cls[u] == k)
                pos = base[k] + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            base[k] += __popcll(bal);
        }
        if (cls[u] >= 0) csorted[pos] = q[u];
    }
}

__device__ __forceinline__ void class_count(const int (&cls)[CLS_UNROLL], int nu, int (&cnt)[8]) {
#pragma unroll
    for (int u = 0; u < CLS_UNROLL; u++) {
        if (u >= nu) break;
#pragma unroll
        for (int k = 0; k < 8; k++) cnt[k] += __popcll(__ballot(cls[u] == k));
    }
}

// Chunks of 64 entries holding any of [j0, end), at most CLS_UNROLL.
__device__ __forceinline__ int class_chunks(int j0, int end) { return min((end - j0 + 63) >> 6, CLS_UNROLL); }

__device__ __forceinline__ uint32_t wave_or3(uint32_t bb) {
    uint32_t r = 0;
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (__ballot((bb >> a) & 1u)) r |= 1u << a;
    return r;
}

// One wave per bin.  A bin of at most CLS_SPAN entries is counted and
// placed from registers.  A larger bin is cut into units of CLS_SPAN entries
// (in order), appended to units / bigs (ctr[0] units, ctr[1] bins, zero on
// entry) for k_class_count, k_class_scan and k_class_place.
constexpr int CLS_SPAN = 64 * CLS_UNROLL;

__global__ __launch_bounds__(64) void k_gather_classes(const float4* sorted, const int32_t* off, CellGrid gb, int logr0,
                                                       float R, int sc, int scs, float cut, Quant16 qz,
                                                       float4* csorted, int32_t* coff, uint8_t* both, int32_t* ctr,
                                                       int2* units, int4* bigs) {
    const int lane = threadIdx.x;
    const int64_t nbin = gb.count();
    __shared__ int32_t lcnt[8];
    for (int64_t ci = blockIdx.x; ci < nbin; ci += gridDim.x) {
        const int beg = off[ci], end = off[ci + 1];
        if (end - beg > CLS_SPAN) {
            const int nu = (end - beg + CLS_SPAN - 1) / CLS_SPAN;
            int ub = 0;
            if (lane == 0) {
                ub = atomicAdd(&ctr[0], nu);
                bigs[atomicAdd(&ctr[1], 1)] = make_int4((int)ci, ub, nu, 0);
            }
            ub = __shfl(ub, 0);
            for (int u = lane; u < nu; u += 64) units[ub + u] = make_int2((int)ci, u);
            continue;
        }
        float4 q[CLS_UNROLL];
        int cls[CLS_UNROLL];
        uint32_t bb = 0;
        int base[8];
#pragma unroll
        for (int k = 0; k < 8; k++) base[k] = 0;
        // counted and placed with LDS atomics: the order inside a class is
        // not kept (the walks take minima, which no order changes)
        const int nu = class_chunks(beg, end);
        if (end > beg) {
            class_load(sorted, beg, end, logr0, R, sc, scs, cut, qz, q, cls, bb);
            if (lane < 8) lcnt[lane] = 0;
            __syncthreads();
#pragma unroll
            for (int u = 0; u < CLS_UNROLL; u++)
                if (u < nu && cls[u] >= 0) atomicAdd(&lcnt[cls[u]], 1);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 8; k++) base[k] = lcnt[k];
            __syncthreads();
        }
        int acc = beg;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int c = base[k];
            base[k] = acc;
            if (lane == k) coff[ci * 8 + k] = acc;
            acc += c;
        }
        const uint32_t bb_all = wave_or3(bb);
        if (lane == 0) both[ci] = (uint8_t)bb_all;
        if (end > beg) {
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < 8; k++) lcnt[k] = base[k];
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < CLS_UNROLL; u++)
                if (u < nu && cls[u] >= 0) csorted[atomicAdd(&lcnt[cls[u]], 1)] = q[u];
            __syncthreads();
        }
    }
    if (blockIdx.x == 0 && lane == 0) coff[nbin * 8] = off[nbin];
}

// Entries of unit ui: its bin and span.
__device__ __forceinline__ int64_t unit_span(const int2* units, const int32_t* off, int ui, int& beg, int& end) {
    const int2 u = units[ui];
    const int b0 = off[u.x];
    beg = b0 + u.y * CLS_SPAN;
    end = min(off[u.x + 1], beg + CLS_SPAN);
    return u.x;
}

// Per unit (one wave each): class counts to ucnt[9 ui + k], both-face axes
// to ucnt[9 ui + 8].
__global__ __launch_bounds__(64) void k_class_count(const float4* sorted, const int32_t* off, int logr0,
                                                    float R, int sc, int scs, float cut, Quant16 qz, const int32_t* ctr,
                                                    const int2* units, int32_t* ucnt) {
    const int lane = threadIdx.x;
    const int nunit = ctr[0];
    for (int ui = blockIdx.x; ui < nunit; ui += gridDim.x) {
        int beg, end;
        unit_span(units, off, ui, beg, end);
        float4 q[CLS_UNROLL];
        int cls[CLS_UNROLL];
        uint32_t bb = 0;
        class_load(sorted, beg, end, logr0, R, sc, scs, cut, qz, q, cls, bb);
        int cnt[8];
#pragma unroll
        for (int k = 0; k < 8; k++) cnt[k] = 0;
        class_count(cls, class_chunks(beg, end), cnt);
        const uint32_t bbw = wave_or3(bb);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (lane == k) ucnt[(int64_t)ui * 9 + k] = cnt[k];
        if (lane == 8) ucnt[(int64_t)ui * 9 + 8] = (int32_t)bbw;
    }
}

// Per large bin (one wave each): class starts (coff), both-face axes, and
// each unit's first position per class, written over its counts.
__global__ __launch_bounds__(64) void k_class_scan(const int32_t* off, const int32_t* ctr, const int4* bigs,
                                                   int32_t* ucnt, int32_t* coff, uint8_t* both) {
    const int lane = threadIdx.x;
    const int nbig = ctr[1];
    for (int bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const int4 bg = bigs[bi];
        const int64_t ci = bg.x;
        const int ub = bg.y, nu = bg.z;
        int tot[8];
        uint32_t bbo = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) tot[k] = 0;
        for (int u0 = 0; u0 < nu; u0 += 64) {
            const int u = u0 + lane;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                int c = u < nu ? ucnt[(int64_t)(ub + u) * 9 + k] : 0;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
                tot[k] += c;
            }
            bbo |= u < nu ? (uint32_t)ucnt[(int64_t)(ub + u) * 9 + 8] : 0u;
        }
        bbo = wave_or3(bbo);
        int run[8];
        int acc = off[ci];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            run[k] = acc;
            if (lane == k) coff[ci * 8 + k] = acc;
            acc += tot[k];
        }
        if (lane == 0) both[ci] = (uint8_t)bbo;
        for (int u0 = 0; u0 < nu; u0 += 64) {
            const int u = u0 + lane;
            int c[8];
#pragma unroll
            for (int k = 0; k < 8; k++) c[k] = u < nu ? ucnt[(int64_t)(ub + u) * 9 + k] : 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                int incl = c[k];   // inclusive scan over the lanes
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(incl, o);
                    if (lane >= o) incl += t;
                }
                if (u < nu) ucnt[(int64_t)(ub + u) * 9 + k] = run[k] + incl - c[k];
                run[k] += __shfl(incl, 63);
            }
        }
    }
}

// Per unit (one wave each): its entries placed from the positions of k_class_scan.
__global__ __launch_bounds__(64) void k_class_place(const float4* sorted, const int32_t* off, int logr0,
                                                    float R, int sc, int scs, float cut, Quant16 qz, const int32_t* ctr,
                                                    const int2* units, const int32_t* ucnt, float4* csorted) {
    const int nunit = ctr[0];
    for (int ui = blockIdx.x; ui < nunit; ui += gridDim.x) {
        int beg, end;
        unit_span(units, off, ui, beg, end);
        float4 q[CLS_UNROLL];
        int cls[CLS_UNROLL];
        uint32_t bb = 0;
        class_load(sorted, beg, end, logr0, R, sc, scs, cut, qz, q, cls, bb);
        int base[8];
#pragma unroll
        for (int k = 0; k < 8; k++) base[k] = ucnt[(int64_t)ui * 9 + k];
        class_place(q, cls, class_chunks(beg, end), base, csorted);
    }
}

struct ChunkIter {
    int d = -1;       // position in NB_ORDER
    int j0 = 0, end = 0;
};

// Neighbour bins nearest first: the tile's own bin, the 6 faces, the 12
// edges, the 8 corners (index (dx+1)*9 + (dy+1)*3 + (dz+1)).  Near points
// come first, so the running minima (and the pruning bound) tighten early.
__constant__ int8_t NB_ORDER[27] = {13, 4, 10, 12, 14, 16, 22, 1, 3, 5, 7, 9, 11, 15,
                                    17, 19, 21, 23, 25, 0, 2, 6, 8, 18, 20, 24, 26};

// Advance k >= 1 chunks through the candidate sequence of the tile in cell
// (cx,cy,cz).  From the initial state (d = -1), k = s + 1 lands on chunk s.
__device__ __forceinline__ bool skip_chunks(ChunkIter& it, const GatherParams& gp, int cx, int cy, int cz, int k) {
    if (it.d >= 0) {
        const int left = (it.end - it.j0 - 1) >> 6;   // chunks after the current one in its bin
        if (k <= left) {
            it.j0 += 64 * k;
            return true;
        }
        k -= left;   // step `left + 1` is the next bin's first chunk
    }
    k -= 1;
    while (++it.d < 27) {
        const int nb = NB_ORDER[it.d];
        int bx = cx + nb / 9 - 1, by = cy + (nb / 3) % 3 - 1, bz = cz + nb % 3 - 1;
        const int64_t ci = gp.gb.find(bx, by, bz);
        if (ci < 0) continue;
        int beg = gp.off[ci], end = gp.off[ci + 1];
        if (beg < end) {
            const int nch = (end - beg + 63) >> 6;
            if (k < nch) {
                it.j0 = beg + 64 * k;
                it.end = end;
                return true;
            }
            k -= nch;
        }
    }
    return false;
}

// Classes of the bin at offset o from the tile's cell that can reach the cell.
__device__ __forceinline__ uint32_t class_mask(int ox, int oy, int oz, uint32_t both) {
    const int o[3] = {ox, oy, oz};
    const uint32_t up[3] = {0xAAu, 0xCCu, 0xF0u};   // classes with bit a set
    uint32_t m = 0xFFu;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (o[a] < 0)
            m &= up[a];
        else if (o[a] > 0 && !((both >> a) & 1u))
            m &= ~up[a];
    }
    return m & 0xFFu;
}

// A tile's candidates with reach classes: the runs of consecutive wanted
// classes of its 27 neighbour bins (at most 4 per bin), nearest bins first.
// Run r starts at entry start[r] and is preceded by pre[r] candidates;
// pre[nruns .. 127] = INT_MAX, so a 7-step binary search finds the run of
// any candidate number.
struct RunTable {
    int pre[128];
    int start[128];
};

// Builds the table (lane d reads neighbour bin d's class starts) and returns
// the number of candidates.
__device__ __forceinline__ int build_runs(const GatherParams& gp, int cx, int cy, int cz, RunTable& rt) {
    const int lane = threadIdx.x;
    int c[9];
    uint32_t want = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) c[k] = 0;
    if (lane < 27) {
        const int nb = NB_ORDER[lane];
        const int ox = nb / 9 - 1, oy = (nb / 3) % 3 - 1, oz = nb % 3 - 1;
        const int64_t ci = gp.gb.find(cx + ox, cy + oy, cz + oz);
        if (ci >= 0) {
#pragma unroll
            for (int k = 0; k < 9; k++) c[k] = gp.coff[ci * 8 + k];
            if (c[8] > c[0]) want = class_mask(ox, oy, oz, gp.cboth[ci]);
        }
    }
    // this lane's non-empty runs: count and entries, then (second pass) stores
    auto runs = [&](auto&& emit) {
        bool in = false;
        int rb = 0;
#pragma unroll
        for (int k = 0; k <= 8; k++) {
            const bool w = k < 8 && ((want >> k) & 1u);
            if (w && !in) rb = c[k];
            if (!w && in && c[k] > rb) emit(rb, c[k] - rb);
            in = w;
        }
    };
    int nr = 0, tot = 0;
    runs([&](int, int len) {
        nr++;
        tot += len;
    });
    int inr = nr, itot = tot;   // inclusive scans over the lanes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int a = __shfl_up(inr, o), b = __shfl_up(itot, o);
        if (lane >= o) {
            inr += a;
            itot += b;
        }
    }
    const int nruns = __builtin_amdgcn_readlane(inr, 63), total = __builtin_amdgcn_readlane(itot, 63);
    int r = inr - nr, before = itot - tot;
    runs([&](int beg, int len) {
        rt.start[r] = beg;
        rt.pre[r] = before;
        r++;
        before += len;
    });
    for (int q = nruns + lane; q < 128; q += 64) rt.pre[q] = INT_MAX;
    __syncthreads();
    return total;
}

// Entry of candidate v (0 <= v < the table's total).
__device__ __forceinline__ int run_entry(const RunTable& rt, int v) {
    int r = 0;
#pragma unroll
    for (int st = 64; st >= 1; st >>= 1)
        if (rt.pre[r + st] <= v) r += st;
    return rt.start[r] + (v - rt.pre[r]);
}

// Where a tile sits: its leaf k, its origin inside the brick, its first
// voxel in (cloud-local) index space and its level-0 cell.
struct GTile {
    int64_t k;
    int tox, toy, toz;
    int T0x, T0y, T0z;
    int cx, cy, cz;
};

__device__ __forceinline__ GTile gather_tile(const GatherParams& gp, int res0, int tiles_per_brick, int64_t tile) {
    GTile g;
    g.k = tile / tiles_per_brick;
    tile_origin(res0, tile % tiles_per_brick, g.tox, g.toy, g.toz);
    const GNode nd = gp.leaves[g.k];
    int shx = 0, shy = 0, shz = 0;   // cloud-local voxel coordinates (batched clouds)
    if (gp.leaf_cloud) {
        const int cl = gp.leaf_cloud[g.k];
        shx = gp.cell_off[3 * cl] * res0, shy = gp.cell_off[3 * cl + 1] * res0, shz = gp.cell_off[3 * cl + 2] * res0;
    }
    g.T0x = nd.mPos[0] - shx + g.tox, g.T0y = nd.mPos[1] - shy + g.toy, g.T0z = nd.mPos[2] - shz + g.toz;
    g.cx = nd.mPos[0] / res0, g.cy = nd.mPos[1] / res0, g.cz = nd.mPos[2] / res0;
    return g;
}

// Wave-wide max of an unsigned value, returned as a scalar: DPP steps give
// every lane its row's max (quad swaps, half-row and row mirrors), two row
// broadcasts carry the row maxima into row 3, and lane 63 holds the result.
// Needs all 64 lanes active.  Lanes a DPP step does not write get 0, the
// identity of an unsigned max.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));    // quad_perm [1,0,3,2]
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));    // quad_perm [2,3,0,1]
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));   // row_half_mirror
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));   // row_mirror
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));   // row_bcast15
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));   // row_bcast31
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Bit ly * td + lx for every column (lx, ly) of the tile with lx in [xlo, xhi]
// and ly in [ylo, yhi] (non-empty ranges inside the tile).
__device__ __forceinline__ uint64_t column_mask(int xlo, int xhi, int ylo, int yhi, int td) {
    const uint64_t xb = ((2ull << xhi) - 1ull) & ~((1ull << xlo) - 1ull);
    if (td == 8) {
        // one byte per row: replicate the row bits, keep rows ylo .. yhi
        const uint64_t rows = ((2ull << (8 * yhi + 7)) - 1ull) & ~((1ull << (8 * ylo)) - 1ull);
        return (xb * 0x0101010101010101ull) & rows;
    }
    uint64_t m = 0;
    for (int y = ylo; y <= yhi; y++) m |= xb << (y * td);
    return m;
}

// Folds chunks s, s + ns, ... of the tile's candidates into m (lane = one
// (x, y) column of the tile, m[z] = bits of min |p-v|^2 over its z voxels;
// non-negative floats order like their bit patterns).  ns = plan(ns) is
// taken after the first chunk's load is issued, so work that decides the
// stride overlaps that load.  CLS: the candidates are the class runs that
// can reach the tile (k_gather_classes), packed into full chunks; otherwise
// every point of the 27 bins, bin by bin.
template <bool HALVES, bool CLS, typename Plan, bool QUART = false>
__device__ __forceinline__ void gather_walk(const GatherParams& gp, const GTile& g, int td_rt, int s, int ns,
                                            Plan&& plan, uint32_t (&m)[8]) {
    const int td = HALVES ? 8 : td_rt;   // HALVES: 8-voxel tiles
    const int lane = threadIdx.x;
    const int T0x = g.T0x, T0y = g.T0y, T0z = g.T0z;
    const int cx = g.cx, cy = g.cy, cz = g.cz;
    const int lx = lane % td, ly = lane / td;
    const bool lane_ok = ly < td;
    const float vx = (float)(T0x + lx), vy = (float)(T0y + ly);
    float vz[8];
#pragma unroll
    for (int z = 0; z < 8; z++) vz[z] = (float)(T0z + z);
    const float R = gp.box_radius;
    const int sc = gp.sc, scs = gp.sc_shift;
    const float cut = gp.cutoff;
    const float cut2 = cut * cut;

    ChunkIter it;
    __shared__ RunTable rt;   // CLS only
    int total = 0, v = 0;     // CLS: candidates of the tile, this lane's candidate number
    bool have;
    if constexpr (CLS) {
        total = build_runs(gp, cx, cy, cz, rt);
        v = 64 * s + lane;
        have = 64 * s < total;
    } else {
        have = skip_chunks(it, gp, cx, cy, cz, s + 1);
    }
    float4 cur = make_float4(0, 0, 0, 0);
    uint32_t curb = 0;   // packed ranges (gp.box)
    if (have && (CLS ? v < total : it.j0 + lane < it.end)) {
        const int j = CLS ? run_entry(rt, v) : it.j0 + lane;
        cur = gp.sorted[j];
        if (CLS)
            curb = __float_as_uint(cur.w);
        else if (gp.box)
            curb = gp.box[j];
    }
    ns = plan(ns, total);
    // One loop for ns == 1 (the stride folds away) and one for split tiles.
    auto walk = [&](auto one) {
        while (have) {
            const bool cvalid = CLS ? v < total : it.j0 + lane < it.end;
            if constexpr (CLS) {
                v += 64 * (decltype(one)::value ? 1 : ns);
                have = v - lane < total;
            } else {
                have = skip_chunks(it, gp, cx, cy, cz, decltype(one)::value ? 1 : ns);
            }
            float4 nxt = make_float4(0, 0, 0, 0);
            uint32_t nxtb = 0;
            if (have && (CLS ? v < total : it.j0 + lane < it.end)) {
                const int j = CLS ? run_entry(rt, v) : it.j0 + lane;
                nxt = gp.sorted[j];
                if (CLS)
                    nxtb = __float_as_uint(nxt.w);
                else if (gp.box)
                    nxtb = gp.box[j];
            }
            // Pruning bound: every voxel's running min is <= the tile's max of
            // them, and a candidate at squared distance db2 from the voxels it
            // may reach changes none of them if db2 exceeds that max (fma is
            // monotone, so every reached voxel's |p-v|^2 >= db2).  Same per lane.
            uint32_t lmax = 0;
            float qbound[4];   // HALVES && QUART: the bound of each 2-voxel z quarter alone
            uint32_t wmax;
            if constexpr (HALVES && QUART) {
                // per quarter: the max over the tile of its voxels' minima (a
                // quarter far from the tile's worst voxels prunes harder)
                uint32_t qm[4];
#pragma unroll
                for (int q = 0; q < 4; q++) qm[q] = lane_ok ? max(m[2 * q], m[2 * q + 1]) : 0u;
                lmax = max(max(qm[0], qm[1]), max(qm[2], qm[3]));
                wmax = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t wq = wave_max_u32(qm[q]);
                    wmax = max(wmax, wq);
                    qbound[q] = fminf(cut2, __uint_as_float(wq));
                }
            } else {
                if (lane_ok) {
#pragma unroll
                    for (int z = 0; z < 8; z++)
                        if (z < td) lmax = max(lmax, m[z]);
                }
                wmax = wave_max_u32(lmax);
            }
            const float bound = fminf(cut2, __uint_as_float(wmax));
            const float lbound = fminf(cut2, __uint_as_float(lmax));
            // Below a bound of R^2 the column masks are implied: a column
            // outside a candidate's range along x or y is a non-member, more
            // than R from it along that axis (its |p-v|^2 >= fl(R*R) > every
            // lane bound, so the lane test of the walk drops it), or beyond
            // the cutoff (which leaves the stored value as it is).
            const bool use_cols = !(bound < R * R);
            // Filter this chunk.  Membership = subcell box (cuda_gvdb_particles.cuh:237-242)
            // as a voxel range; a candidate must also come within the cutoff of the tile.
            // The packed range is the membership range clipped to the tile (z) and to
            // membership + cutoff (x, y: lanes outside it are masked).
            bool ok = cvalid;
            uint32_t pk = 0;   // z range inside the tile: lo | hi << 3
            uint64_t cm = 0;   // tile columns in the x/y range: bit ly * td + lx (= their lane)
            {
                float pw[3] = {cur.x, cur.y, cur.z};
                int T0[3] = {T0x, T0y, T0z};
                int rlo[2] = {0, 0}, rhi[2] = {0, 0};
                float db2 = 0.0f, db2xy = 0.0f;
                if (CLS || gp.box) ok = ok && !(curb >> 31);
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    int lo = 0, hi = 0, lo_c, hi_c;
                    if (CLS || gp.box) {
                        // the ranges from the pre-pass; z of the halves walk then
                        // uses the clipped range too (voxels beyond the cutoff
                        // cannot change the result)
                        const int f = (int)floorf(pw[a]);
                        lo_c = f + 1 - (int)((curb >> (10 * a)) & 31u);
                        hi_c = f + (int)((curb >> (10 * a + 5)) & 31u);
                        lo = lo_c, hi = hi_c;
                    } else {
                        lo = gvdbx_cdiv_mul((int)(pw[a] - R), scs);
                        hi = gvdbx_cdiv_mul((int)(pw[a] + 1 + R), scs) + sc - 1;
                        lo_c = max(lo, (int)ceilf(pw[a] - cut));
                        hi_c = min(hi, (int)floorf(pw[a] + cut));
                    }
                    int lr = max(lo_c - T0[a], 0), hr = min(hi_c - T0[a], td - 1);
                    ok = ok && lr <= hr;
                    float e = fmaxf(fmaxf((float)(T0[a] + lr) - pw[a], pw[a] - (float)(T0[a] + hr)), 0.0f);
                    if (a == 2) db2xy = db2;
                    db2 = __builtin_fmaf(e, e, db2);
                    if (a < 2) {
                        rlo[a] = lr, rhi[a] = hr;
                    } else {
                        if (HALVES && QUART) {
                            // z: membership (whole 4-voxel halves, sc >= 4), then
                            // per 2-voxel quarter the test below for halves
                            lr = max(lo - T0[a], 0);
                            hr = min(hi - T0[a], td - 1);
                            uint32_t qm = 0;
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                const float eq = fmaxf(fmaxf((float)(T0[a] + 2 * q) - pw[a], pw[a] - (float)(T0[a] + 2 * q + 1)), 0.0f);
                                const bool in = 2 * q + 1 >= lr && 2 * q <= hr && !(__builtin_fmaf(eq, eq, db2xy) > qbound[q]);
                                qm |= in ? 1u << q : 0u;
                            }
                            ok = ok && qm != 0;
                            pk = qm;
                        } else if (HALVES) {
                            // z: membership only (voxels beyond the cutoff cannot change the result)
                            lr = max(lo - T0[a], 0);
                            hr = min(hi - T0[a], td - 1);
                            // A z-half whose nearest voxel is farther than the bound
                            // (x and y at their nearest columns) holds no voxel the
                            // candidate can lower: fma is monotone, so every voxel
                            // there has |p-v|^2 >= that distance > bound >= its
                            // running minimum (or > cutoff^2).  Same arithmetic as db2.
                            const float e0 = fmaxf(fmaxf((float)T0[a] - pw[a], pw[a] - (float)(T0[a] + 3)), 0.0f);
                            const float e1 = fmaxf(fmaxf((float)(T0[a] + 4) - pw[a], pw[a] - (float)(T0[a] + 7)), 0.0f);
                            if (__builtin_fmaf(e0, e0, db2xy) > bound) lr = max(lr, 4);
                            if (__builtin_fmaf(e1, e1, db2xy) > bound) hr = min(hr, 3);
                            ok = ok && lr <= hr;
                        }
                        if (!(HALVES && QUART)) pk = (uint32_t)(lr & 7) | ((uint32_t)(hr & 7) << 3);
                    }
                }
                ok = ok && db2 <= bound;
                if (ok && use_cols) cm = column_mask(rlo[0], rhi[0], rlo[1], rhi[1], td);
            }
            uint64_t mask = __ballot(ok);
            while (mask) {
                const int b = __builtin_ctzll(mask);
                mask &= mask - 1;
                const float px = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.x), b));
                const float py = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.y), b));
                const float pz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.z), b));
                const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)pk, b);
                uint64_t cols = ~0ull;
                if (use_cols)
                    cols = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)cm, b) |
                           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cm >> 32), b) << 32);
                const int zlo = q & 7, zhi = (q >> 3) & 7;
                // the lanes whose column the candidate reaches, straight from the mask
                if (__builtin_amdgcn_inverse_ballot_w64(cols)) {
                    const float dx = vx - px, dy = vy - py;
                    const float dxy = __builtin_fmaf(dy, dy, dx * dx);
                    if (dxy <= lbound) {
                        if (HALVES && QUART) {
                            // q: the 2-voxel z quarters the candidate can lower
#pragma unroll
                            for (int qq = 0; qq < 4; qq++) {
                                if ((q >> qq) & 1u) {
#pragma unroll
                                    for (int z = 2 * qq; z < 2 * qq + 2; z++) {
                                        const float dz = vz[z] - pz;
                                        m[z] = min(m[z], __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
                                    }
                                }
                            }
                        } else if (HALVES) {
                            // with sc >= 4 the z membership inside an 8-voxel tile is
                            // one or both 4-voxel halves
                            if (zlo < 4) {
#pragma unroll
                                for (int z = 0; z < 4; z++) {
                                    const float dz = vz[z] - pz;
                                    m[z] = min(m[z], __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
                                }
                            }
                            if (zhi >= 4) {
#pragma unroll
                                for (int z = 4; z < 8; z++) {
                                    const float dz = vz[z] - pz;
                                    m[z] = min(m[z], __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
                                }
                            }
                        } else {
#pragma unroll
                            for (int z = 0; z < 8; z++) {
                                if (z >= zlo && z <= zhi) {
                                    const float dz = vz[z] - pz;
                                    m[z] = min(m[z], __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
                                }
                            }
                        }
                    }
                }
            }
            cur = nxt;
            curb = nxtb;
        }
    };
    if (ns == 1)
        walk(std::true_type{});
    else
        walk(std::false_type{});
}

// Brick values of a tile from its minima: sqrt is correctly rounded, so the
// stored value is a function of min |p-v|^2 alone.
//
// gp.push (8^3 bricks, apron 1, one tile per brick): the tile's boundary
// voxels are also written into the aprons of the neighbour leaves that hold
// them as apron voxels (faces, edges and corners), the words UpdateApron
// would copy there (neighbours found as k_apron10 finds them).  Every apron
// voxel that has a neighbour leaf then holds that leaf's new value, and
// UpdateApron only has to write the boundary value where no leaf is.  (The
// x faces cost the most: one word per 40-byte row of the neighbour; copying
// them in UpdateApron instead measured no faster.)
//
// push_leaves: lane d < 27 gets the leaf at direction d = (dz+1)*9 + (dy+1)*3
// + (dx+1) from the tile's cell (-1 where none); taken before the walk, so
// only this one register stays live across it.
__device__ __forceinline__ int32_t push_leaves(const GatherParams& gp, const GTile& g) {
    const int lane = threadIdx.x;
    int32_t nl = -1;
    if (gp.push && lane < 27) {
        const int64_t ci = gp.gb.find(g.cx + lane % 3 - 1, g.cy + (lane / 3) % 3 - 1, g.cz + lane / 9 - 1);
        if (ci >= 0) nl = gp.bin_leaf[ci];
    }
    return nl;
}

__device__ __forceinline__ void gather_store(const GatherParams& gp, const GTile& g, int td, const uint32_t (&m)[8],
                                             int32_t nl) {
    const int lane = threadIdx.x;
    const int lx = lane % td, ly = lane / td;
    if (ly >= td) return;
    const int br = gp.brick_res, a = gp.apron;
    float* brick = gp.chan + g.k * br * br * br;
    const float r = gp.radius;
    float val[8];
#pragma unroll
    for (int z = 0; z < 8; z++) {
        if (z < td) {
            int64_t vi = ((int64_t)(g.toz + z + a) * br + (g.toy + ly + a)) * br + (g.tox + lx + a);
            float init = gp.accum ? brick[vi] : (gp.mode == 1 ? 3.0f : 0.0f);
            float c = sqrtf(__uint_as_float(m[z]));
            val[z] = gp.mode == 1 ? fminf(init, c - r) : fmaxf(init, r - c);
            brick[vi] = val[z];
        }
    }
    if (!gp.push) return;
#pragma unroll 1
    for (int d = 0; d < 27; d++) {
        if (d == 13) continue;
        const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
        const int32_t n = __shfl(nl, d);   // every lane takes part (td == 8)
        const bool col = (dx == 0 || lx == (dx < 0 ? 0 : 7)) && (dy == 0 || ly == (dy < 0 ? 0 : 7));
        if (n < 0 || !col) continue;
        float* nb = gp.chan + (int64_t)n * 1000;
        const int px = dx < 0 ? 9 : (dx > 0 ? 0 : lx + 1), py = dy < 0 ? 9 : (dy > 0 ? 0 : ly + 1);
        if (dz == 0) {
#pragma unroll
            for (int z = 0; z < 8; z++) nb[((z + 1) * 10 + py) * 10 + px] = val[z];
        } else {
            nb[((dz < 0 ? 9 : 0) * 10 + py) * 10 + px] = val[dz < 0 ? 0 : 7];
        }
    }
}

// One wave per tile.  SPLIT: a tile with more than 2 split chunks is cut
// into ns = ceil(chunks / split) items; this wave takes a scratch slot and
// item 0, appends items 1 .. ns-1 to `items` (counts[0], zero on entry) for
// k_gather_items, and leaves its minima in scratch tile `slot` (counts[1]
// slots).  Appends beyond item_cap / slot_cap are dropped: the host sizes
// both from a bound that holds.
struct GatherSplit {
    int split;
    int4* items;
    int64_t item_cap;
    int32_t* counts;
    int32_t* slot_tile;
    int64_t slot_cap;
    uint32_t* acc;
};

template <bool HALVES, bool SPLIT, bool CLS, bool QUART = false>
__global__ __launch_bounds__(64) void k_gather_rl(GatherParams gp, int res0, int tiles_per_brick, int64_t nleaf_tiles,
                                                  GatherSplit gs) {
    const int64_t tile = gvdbx_xcd_item(blockIdx.x, nleaf_tiles, gp.xcd_remap);
    if (tile < 0) return;
    if (tile / tiles_per_brick >= gp.nleaf) return;
    const int lane = threadIdx.x;
    const int td = tile_dim(res0);
    const GTile g = gather_tile(gp, res0, tiles_per_brick, tile);
    const int32_t nl = push_leaves(gp, g);
    int ns = 1, slot = -1;
    auto plan = [&](int, int total) {
        if (!SPLIT) return 1;
        int nch = 0;   // lane d < 27: chunks of neighbour bin d (CLS: the run table's total)
        if (CLS) {
            nch = (total + 63) >> 6;
        } else {
            if (lane < 27) {
                const int bx = g.cx + lane / 9 - 1, by = g.cy + (lane / 3) % 3 - 1, bz = g.cz + lane % 3 - 1;
                const int64_t ci = gp.gb.find(bx, by, bz);
                if (ci >= 0) {
                    nch = (gp.off[ci + 1] - gp.off[ci] + 63) >> 6;
                }
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) nch += __shfl_xor(nch, o);
            nch = __builtin_amdgcn_readfirstlane(nch);
        }
        if (nch > 2 * gs.split) {
            ns = (nch + gs.split - 1) / gs.split;
            int base = 0, sl = 0;
            if (lane == 0) {
                base = atomicAdd(&gs.counts[0], ns - 1);
                sl = atomicAdd(&gs.counts[1], 1);
            }
            base = __builtin_amdgcn_readfirstlane(base);
            sl = __builtin_amdgcn_readfirstlane(sl);
            slot = sl < gs.slot_cap ? sl : -1;
            for (int s = 1 + lane; s < ns; s += 64)
                if ((int64_t)base + s - 1 < gs.item_cap) gs.items[base + s - 1] = make_int4((int)tile, s, ns, slot);
            if (lane == 0 && slot >= 0) gs.slot_tile[slot] = (int)tile;
        }
        return ns;
    };
    uint32_t m[8];
#pragma unroll
    for (int z = 0; z < 8; z++) m[z] = 0x7F800000u;
    gather_walk<HALVES, CLS, decltype(plan)&, QUART>(gp, g, td, 0, 1, plan, m);
    if (!SPLIT || ns == 1) {
        gather_store(gp, g, td, m, nl);
    } else if (slot >= 0) {
        const bool lane_ok = lane / td < td;
        uint32_t* at = gs.acc + (int64_t)slot * 512 + lane;
#pragma unroll
        for (int z = 0; z < 8; z++) at[z * 64] = lane_ok && z < td ? m[z] : 0x7F800000u;
    }
}

// Persistent walk over the items beyond the first of each split tile
// (counts[0] of them; the scratch tiles hold item 0's minima).
template <bool HALVES, bool CLS, bool QUART = false>
__global__ __launch_bounds__(64) void k_gather_items(GatherParams gp, int res0, int tiles_per_brick,
                                                     const int4* items, int64_t item_cap, const int32_t* counts,
                                                     uint32_t* acc, int64_t slot_cap) {
    const int lane = threadIdx.x;
    const int td = tile_dim(res0);
    const int nitems = (int)min((int64_t)counts[0], item_cap);
    for (int idx = blockIdx.x; idx < nitems; idx += gridDim.x) {
        const int4 w = items[idx];
        const GTile g = gather_tile(gp, res0, tiles_per_brick, w.x);
        uint32_t m[8];
#pragma unroll
        for (int z = 0; z < 8; z++) m[z] = 0x7F800000u;
        auto same = [](int n, int) { return n; };
        gather_walk<HALVES, CLS, decltype(same)&, QUART>(gp, g, td, w.y, w.z, same, m);
        if (lane / td < td && w.w >= 0 && w.w < slot_cap) {
            uint32_t* at = acc + (int64_t)w.w * 512 + lane;
#pragma unroll
            for (int z = 0; z < 8; z++)
                if (z < td && m[z] != 0x7F800000u) atomicMin(at + z * 64, m[z]);
        }
    }
}

// Bricks of the split tiles from their scratch minima.
__global__ __launch_bounds__(64) void k_gather_finish(GatherParams gp, int res0, int tiles_per_brick,
                                                      const int32_t* slot_tile, const int32_t* counts,
                                                      const uint32_t* acc, int64_t slot_cap) {
    const int lane = threadIdx.x;
    const int td = tile_dim(res0);
    const int nslot = (int)min((int64_t)counts[1], slot_cap);
    for (int sl = blockIdx.x; sl < nslot; sl += gridDim.x) {
        const GTile g = gather_tile(gp, res0, tiles_per_brick, slot_tile[sl]);
        uint32_t m[8];
#pragma unroll
        for (int z = 0; z < 8; z++) m[z] = acc[(int64_t)sl * 512 + z * 64 + lane];
        gather_store(gp, g, td, m, push_leaves(gp, g));
    }
}

// ---------------------------------------------------------------------------
// Weighted cubic splat: rho(v) = sum_p w_p * max(0, 1 - |v-p|^2/r^2)^3 at the
// integer voxel coordinates of every leaf (extension of distFunc,
// cuda_gvdb_particles.cuh:516-523).  Two forms, both deterministic:
//
// Cell push (default).  Pass 1: the points of each bin cell, in runs of at
// most `chunk` points (a tuning knob) per wavefront, are splat into an LDS
// grid covering the cell's voxels plus a halo of h = ceil(r) voxels
// (G = res0 + 2h per axis), which holds every voxel any of them reaches.
// Points are taken one at a time in point order and the lanes cover the
// point's W^3 support window (W = ceil(2r)), so every grid voxel sums its
// contributions in point order with plain LDS adds.  Each point is read once.
// The grid goes to the run's slot of a scratch buffer.  Pass 2, one block per
// leaf: every interior voxel adds, to its initial value, the entries covering
// it of the runs of its own cell and of its 26 neighbour cells, in a fixed
// order (cells nearest first, NB_ORDER; runs in point order).
//
// Tile pull (fallback for grids that do not fit LDS).  One wave per 8^3
// tile scans the 27 neighbour bins nearest first, 64 points per step with the
// next step's load in flight, compacts the points whose support box overlaps
// the tile into LDS in that order, and splats every full round of 64 staged
// points point-centric into the tile's LDS copy.
//
// Tile pull reads the points from an (x, y, z, weight) copy of the bins
// made once per call.  POW2: r^2 is a power of two, so d2 * (1/r^2) equals
// d2 / r^2 bit for bit.
// ---------------------------------------------------------------------------
constexpr int SPLAT_STAGE = 128;   // tile pull: staged points per wave (one round of 64 plus one step)

__global__ __launch_bounds__(256) void k_splat_src(const float4* sorted, int64_t n, const int32_t* count,
                                                   const char* weights, int64_t wstride, float4* out) {
    // Only the first *count entries (the binned points; invalid ones sort
    // last and are never written) are read.
    const int64_t m = min(n, (int64_t)*count);
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
        const float4 p = sorted[j];
        const float w = weights ? *(const float*)(weights + (int64_t)__float_as_int(p.w) * wstride) : 1.0f;
        out[j] = make_float4(p.x, p.y, p.z, w);
    }
}

// Bin of neighbour d (NB_ORDER position) of cell (cx,cy,cz), or -1.
__device__ __forceinline__ int64_t nb_bin(const CellGrid& gb, int d, int cx, int cy, int cz) {
    const int nb = NB_ORDER[d];
    const int bx = cx + nb / 9 - 1, by = cy + (nb / 3) % 3 - 1, bz = cz + nb % 3 - 1;
    return gb.find(bx, by, bz);
}

// ---- cell push ----
__global__ __launch_bounds__(256) void k_splat_runs(const int32_t* off, int64_t ncell, int chunk, int32_t* runs) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int n = off[c + 1] - off[c];
        runs[c] = (int)(((int64_t)n + chunk - 1) / chunk);
    }
}

__global__ __launch_bounds__(256) void k_splat_run_map(const int32_t* run_off, int64_t ncell, int32_t* run_cell) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x)
        for (int32_t b = run_off[c]; b < run_off[c + 1]; b++) run_cell[b] = (int32_t)c;
}

// First voxel of a point's support along one axis, as a grid index: the
// smallest integer x with |x - p| < r (ceil(p - r), one more when that voxel
// is at distance >= r), clamped so that the W-voxel window stays in the grid
// (the voxels the clamp drops are at distance >= r).
__device__ __forceinline__ int support_start(float p, float r, float r2, int g0, int G, int W) {
    int x = (int)ceilf(p - r);
    const float d = (float)x - p;
    if (!(d * d < r2)) x++;
    return min(max(x - g0, 0), G - W);
}

// Pass 1.  The run's points are walked one at a time (ballot + readlane);
// the 64 lanes cover a 4x4x4 block of the point's W^3 support window (several
// blocks when W > 4), so the lanes of one LDS read-add-write never share a
// voxel and need no atomics, and every grid voxel sums its contributions in
// point order.  Weights are fetched two steps ahead.
// SMALL: W <= 4 (r <= 2), a single block.
template <bool POW2, bool SMALL>
__global__ __launch_bounds__(64) void k_splat_cell(SplatParams sp, const int32_t* run_cell, const int32_t* run_off,
                                                   int nruns, int res0, int h, int W, int chunk, float* part) {
    extern __shared__ float grid[];
    const int G = res0 + 2 * h;
    const int G3 = G * G * G;
    const int lane = threadIdx.x;
    const int b = (int)blockIdx.x;
    if (b >= nruns) return;
    const int64_t c = run_cell[b];
    const int sub = b - run_off[c];
    const int beg = sp.off[c] + sub * chunk;
    const int end = (int)min((int64_t)sp.off[c + 1], (int64_t)beg + chunk);
    for (int i = lane; i < G3; i += 64) grid[i] = 0.0f;
    // cell coordinates; batched clouds: shift to the cloud's own voxel space
    int cx, cy, cz;
    sp.gb.cell(c, cx, cy, cz);
    if (sp.cm.ncloud > 0) {
        const int cl = sp.cm.cloud_of_point(__float_as_int(sp.bins[beg].w));
        cx -= sp.cm.cell_off[3 * cl];
        cy -= sp.cm.cell_off[3 * cl + 1];
        cz -= sp.cm.cell_off[3 * cl + 2];
    }
    const int g0x = cx * res0 - h, g0y = cy * res0 - h, g0z = cz * res0 - h;
    const float r = sp.radius;
    const float r2 = r * r;
    const float inv_r2 = 1.0f / r2;
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const float lxf = (float)lx, lyf = (float)ly, lzf = (float)lz;
    const int loff = (lz * G + ly) * G + lx;
    const bool lane_in = lx < W && ly < W && lz < W;   // SMALL: lane inside the window
    auto weight_of = [&](const float4& q, bool valid) -> float {
        if (!valid || !sp.weights) return 1.0f;
        return *(const float*)(sp.weights + (int64_t)__float_as_int(q.w) * sp.wstride);
    };
    auto add = [&](int gi, float d2, float w) {
        // POW2: d2 * inv_r2 is exact, so the fma rounds 1 - d2/r^2 once, as the division form does
        const float t = POW2 ? __builtin_fmaf(-d2, inv_r2, 1.0f) : 1.0f - d2 / r2;
        // a plain read-add-write: the wave waits for its own read, which in
        // dense cells is cheaper than back-to-back LDS atomics on one address
        grid[gi] = grid[gi] + w * ((t * t) * t);
    };
    __syncthreads();
    // pipeline: cur (+ its weight) is splat while the weight of nxt and the
    // points of the step after are in flight
    float4 cur = make_float4(0, 0, 0, 0), nxt = make_float4(0, 0, 0, 0);
    if (beg + lane < end) cur = sp.bins[beg + lane];
    float cw = weight_of(cur, beg + lane < end);
    if (beg + 64 + lane < end) nxt = sp.bins[beg + 64 + lane];
    for (int j0 = beg; j0 < end; j0 += 64) {
        const bool ok = j0 + lane < end;
        const float nw = weight_of(nxt, j0 + 64 + lane < end);
        float4 nn = make_float4(0, 0, 0, 0);
        if (j0 + 128 + lane < end) nn = sp.bins[j0 + 128 + lane];
        // per-lane point: grid index of its support window's first voxel and
        // that voxel's offset from the point
        int pk = 0;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (ok) {
            const int sx = support_start(cur.x, r, r2, g0x, G, W);
            const int sy = support_start(cur.y, r, r2, g0y, G, W);
            const int sz = support_start(cur.z, r, r2, g0z, G, W);
            ax = (float)(g0x + sx) - cur.x;
            ay = (float)(g0y + sy) - cur.y;
            az = (float)(g0z + sz) - cur.z;
            pk = (sz * G + sy) * G + sx;
        }
        uint64_t m = __ballot(ok);
        while (m) {
            const int q = __builtin_ctzll(m);
            m &= m - 1;
            const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ax), q));
            const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ay), q));
            const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(az), q));
            const float qw = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cw), q));
            const int qp = __builtin_amdgcn_readlane(pk, q);
            if (SMALL) {
                const float dx = qx + lxf, dy = qy + lyf, dz = qz + lzf;
                const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                if (lane_in && d2 < r2) add(qp + loff, d2, qw);
            } else {
                for (int bz = 0; bz < W; bz += 4)
                    for (int by = 0; by < W; by += 4)
                        for (int bx = 0; bx < W; bx += 4) {
                            if (bx + lx >= W || by + ly >= W || bz + lz >= W) continue;
                            const float dx = qx + (float)bx + lxf, dy = qy + (float)by + lyf,
                                        dz = qz + (float)bz + lzf;
                            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                            if (d2 < r2) add(qp + loff + (bz * G + by) * G + bx, d2, qw);
                        }
            }
        }
        cur = nxt;
        cw = nw;
        nxt = nn;
    }
    __syncthreads();
    float* out = part + (int64_t)b * G3;
    for (int i = lane; i < G3; i += 64) out[i] = grid[i];
}

// Pass 2: interior voxel (x,y,z) of leaf k = initial value + the entries at
// that voxel of every run of the 27 cells around the leaf's cell, cells in
// NB_ORDER, runs in point order.  A voxel is covered by neighbour (ox,oy,oz)
// only within h of the matching faces; per-axis masks over NB_ORDER
// positions select those neighbours.  R0: res0 at compile time (0: runtime).
template <int R0>
__global__ __launch_bounds__(256) void k_splat_collect(SplatParams sp, const int32_t* run_off, const float* part,
                                                       int res0_rt, int h) {
    const int64_t k = blockIdx.x;
    if (k >= sp.nleaf) return;
    const int res0 = R0 > 0 ? R0 : res0_rt;
    __shared__ int32_t nrun[27];
    __shared__ int32_t goff[27];       // grid index offset of the neighbour's grid
    __shared__ const float* pbase[27];  // first run's grid of the neighbour
    __shared__ uint32_t axm[3][4];      // [axis][low | high << 1]: NB positions whose offset on the axis is allowed
    __shared__ uint32_t nonempty;
    const int G = res0 + 2 * h;
    const int64_t G3 = (int64_t)G * G * G;
    const GNode nd = sp.leaves[k];
    const int cx = nd.mPos[0] / res0, cy = nd.mPos[1] / res0, cz = nd.mPos[2] / res0;
    if (threadIdx.x < 27) {
        const int d = threadIdx.x;
        const int64_t ci = nb_bin(sp.gb, d, cx, cy, cz);
        const int rb = ci >= 0 ? run_off[ci] : 0, re = ci >= 0 ? run_off[ci + 1] : 0;
        nrun[d] = re - rb;
        pbase[d] = part + (int64_t)rb * G3;
        const int nb = NB_ORDER[d];
        const int ox = nb / 9 - 1, oy = (nb / 3) % 3 - 1, oz = nb % 3 - 1;
        goff[d] = ((h - oz * res0) * G + (h - oy * res0)) * G + (h - ox * res0);
    } else if (threadIdx.x >= 32 && threadIdx.x < 44) {
        const int ax = (threadIdx.x - 32) >> 2, cs = (threadIdx.x - 32) & 3;
        uint32_t mm = 0;
        for (int d = 0; d < 27; d++) {
            const int nb = NB_ORDER[d];
            const int o = (ax == 0 ? nb / 9 : ax == 1 ? (nb / 3) % 3 : nb % 3) - 1;
            if (o == 0 || (o < 0 && (cs & 1)) || (o > 0 && (cs & 2))) mm |= 1u << d;
        }
        axm[ax][cs] = mm;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const bool ne = threadIdx.x < 27 && nrun[threadIdx.x] > 0;
        const uint64_t bal = __ballot(ne);
        if (threadIdx.x == 0) nonempty = (uint32_t)bal;
    }
    __syncthreads();
    const uint32_t nem = nonempty;
    const int br = sp.brick_res, a = sp.apron;
    float* brick = sp.chan + k * (int64_t)br * br * br;
    const int nv = res0 * res0 * res0;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        const int x = i % res0, y = (i / res0) % res0, z = i / (res0 * res0);
        const int64_t vi = ((int64_t)(z + a) * br + (y + a)) * br + (x + a);
        float v = sp.accum ? brick[vi] : 0.0f;
        const int vg = (z * G + y) * G + x;
        uint32_t mm = nem & axm[0][(x < h) | ((x >= res0 - h) << 1)] & axm[1][(y < h) | ((y >= res0 - h) << 1)] &
                      axm[2][(z < h) | ((z >= res0 - h) << 1)];
        while (mm) {
            const int d = __builtin_ctz(mm);
            mm &= mm - 1;
            const float* src = pbase[d] + (vg + goff[d]);
            const int n = nrun[d];
            v += src[0];
            for (int q = 1; q < n; q++) v += src[(int64_t)q * G3];
        }
        brick[vi] = v;
    }
}

// Adjoint of the cubic splat over the same runs:
//   dL/dw = sum_v g(v) t^3,  dL/dp = sum_v g(v) w 6 t^2 (v - p) / r^2,
//   t = 1 - |v-p|^2/r^2, over integer voxels v of active leaves, |v-p| < r.
// One wave per run: the upstream gradients of the run's cell plus its halo
// are staged in an LDS grid (0 where no leaf), one leaf lookup per
// neighbour cell; then each lane takes one point and sums over its W^3
// support window, reading LDS only.
template <bool POW2, int WT>
__global__ __launch_bounds__(64) void k_splat_cell_backward(SplatBackParams p, CellGrid gb, const int32_t* off,
                                                            const int32_t* run_cell, const int32_t* run_off, int nruns,
                                                            int res0, int h, int Wrt, int chunk) {
    extern __shared__ float grid[];
    __shared__ int32_t nbleaf[27];   // leaf of neighbour (dx,dy,dz) at dx*9 + dy*3 + dz (offsets + 1)
    const int W = WT > 0 ? WT : Wrt;
    const int G = res0 + 2 * h;
    const int G3 = G * G * G;
    const int lane = threadIdx.x;
    const int b = (int)blockIdx.x;
    if (b >= nruns) return;
    const int64_t c = run_cell[b];
    const int sub = b - run_off[c];
    const int beg = off[c] + sub * chunk;
    const int end = (int)min((int64_t)off[c + 1], (int64_t)beg + chunk);
    int gcx, gcy, gcz;
    gb.cell(c, gcx, gcy, gcz);
    if (lane < 27) {
        const int nx = gcx + lane / 9 - 1, ny = gcy + (lane / 3) % 3 - 1, nz = gcz + lane % 3 - 1;
        int32_t leaf = -1;
        const int64_t ncell = p.g0.find(nx, ny, nz);
        if (ncell >= 0) leaf = p.idx0[ncell];
        nbleaf[lane] = (leaf >= 0 && leaf < p.nleaf) ? leaf : -1;
    }
    __syncthreads();
    const int br = p.brick_res, a = (br - res0) / 2;
    const int64_t nv = (int64_t)br * br * br;
    for (int i = lane; i < G3; i += 64) {
        const int ox = i % G - h, oy = (i / G) % G - h, oz = i / (G * G) - h;   // voxel offset from the cell origin
        const int dx = ox < 0 ? 0 : (ox >= res0 ? 2 : 1), dy = oy < 0 ? 0 : (oy >= res0 ? 2 : 1),
                  dz = oz < 0 ? 0 : (oz >= res0 ? 2 : 1);
        const int32_t leaf = nbleaf[dx * 9 + dy * 3 + dz];
        float v = 0.0f;
        if (leaf >= 0) {
            const int lx = ox - (dx - 1) * res0 + a, ly = oy - (dy - 1) * res0 + a, lz = oz - (dz - 1) * res0 + a;
            v = p.grad[leaf * nv + ((int64_t)lz * br + ly) * br + lx];
        }
        grid[i] = v;
    }
    // positions are in the cloud's own voxel space
    int cx = gcx, cy = gcy, cz = gcz;
    if (p.cm.ncloud > 0) {
        const int cl = p.cm.cloud_of_point(__float_as_int(p.sorted[beg].w));
        cx -= p.cm.cell_off[3 * cl];
        cy -= p.cm.cell_off[3 * cl + 1];
        cz -= p.cm.cell_off[3 * cl + 2];
    }
    const int g0x = cx * res0 - h, g0y = cy * res0 - h, g0z = cz * res0 - h;
    const float r = p.radius;
    const float r2 = r * r;
    const float inv_r2 = 1.0f / r2;
    __syncthreads();
    for (int j = beg + lane; j < end; j += 64) {
        const float4 q = p.sorted[j];
        const int i = __float_as_int(q.w);
        const float wt = p.weights ? *(const float*)(p.weights + (int64_t)i * p.wstride) : 1.0f;
        const int sx = support_start(q.x, r, r2, g0x, G, W);
        const int sy = support_start(q.y, r, r2, g0y, G, W);
        const int sz = support_start(q.z, r, r2, g0z, G, W);
        float dzs[WT > 0 ? WT : 16];
#pragma unroll
        for (int k = 0; k < (WT > 0 ? WT : 16); k++)
            if (k < W) dzs[k] = (float)(g0z + sz + k) - q.z;
        float gw = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
        for (int ix = 0; ix < W; ix++) {
            const float dx = (float)(g0x + sx + ix) - q.x;
            const float dxx = dx * dx;
            for (int iy = 0; iy < W; iy++) {
                const float dy = (float)(g0y + sy + iy) - q.y;
                const float s2 = __builtin_fmaf(dy, dy, dxx);
                const float* row = grid + ((sz * G) + (sy + iy)) * G + (sx + ix);
#pragma unroll
                for (int k = 0; k < (WT > 0 ? WT : 16); k++) {
                    if (k < W) {
                        const float dz = dzs[k];
                        const float d2 = __builtin_fmaf(dz, dz, s2);
                        if (d2 < r2) {
                            const float g = row[k * G * G];
                            const float t = POW2 ? __builtin_fmaf(-d2, inv_r2, 1.0f) : 1.0f - d2 / r2;
                            gw += g * ((t * t) * t);
                            const float f = POW2 ? g * wt * 6.0f * (t * t) * inv_r2 : g * wt * 6.0f * (t * t) / r2;
                            gx += f * dx;
                            gy += f * dy;
                            gz += f * dz;
                        }
                    }
                }
            }
        }
        if (p.dw) p.dw[i] = gw;
        p.dpos[3 * (int64_t)i] = gx;
        p.dpos[3 * (int64_t)i + 1] = gy;
        p.dpos[3 * (int64_t)i + 2] = gz;
    }
}

// ---- tile pull ----
template <class P>
__device__ __forceinline__ bool next_bin_chunk(ChunkIter& it, const P& gp, int cx, int cy, int cz) {
    if (it.j0 + 64 < it.end) {
        it.j0 += 64;
        return true;
    }
    while (++it.d < 27) {
        const int64_t ci = nb_bin(gp.gb, it.d, cx, cy, cz);
        if (ci < 0) continue;
        const int beg = gp.off[ci], end = gp.off[ci + 1];
        if (beg < end) {
            it.j0 = beg;
            it.end = end;
            return true;
        }
    }
    return false;
}

template <bool POW2>
__global__ __launch_bounds__(64) void k_splat_tile(SplatParams sp, int res0, int tiles_per_brick, int64_t nleaf_tiles) {
    __shared__ float4 cand[SPLAT_STAGE];
    __shared__ float acc[512];
    const int lane = threadIdx.x;
    const int64_t tile = gvdbx_xcd_item(blockIdx.x, nleaf_tiles);
    if (tile < 0) return;
    const int64_t k = tile / tiles_per_brick;
    if (k >= sp.nleaf) return;
    const int td = tile_dim(res0);
    const int nv = td * td * td;
    int tox, toy, toz;
    tile_origin(res0, tile % tiles_per_brick, tox, toy, toz);
    const GNode nd = sp.leaves[k];
    int shx = 0, shy = 0, shz = 0;   // cloud-local voxel coordinates (batched clouds)
    if (sp.leaf_cloud) {
        const int cl = sp.leaf_cloud[k];
        shx = sp.cell_off[3 * cl] * res0, shy = sp.cell_off[3 * cl + 1] * res0, shz = sp.cell_off[3 * cl + 2] * res0;
    }
    const int T0x = nd.mPos[0] - shx + tox, T0y = nd.mPos[1] - shy + toy, T0z = nd.mPos[2] - shz + toz;
    const int br = sp.brick_res, a = sp.apron;
    float* brick = sp.chan + (int64_t)k * br * br * br;
    for (int i = lane; i < nv; i += 64) {
        const int x = i % td, y = (i / td) % td, z = i / (td * td);
        acc[i] = sp.accum ? brick[((int64_t)(toz + z + a) * br + (toy + y + a)) * br + (tox + x + a)] : 0.0f;
    }
    const int cx = nd.mPos[0] / res0, cy = nd.mPos[1] / res0, cz = nd.mPos[2] / res0;
    const float r = sp.radius;
    const float r2 = r * r;
    const float inv_r2 = 1.0f / r2;

    // lane `lane` splats staged point `slot` (slot < count)
    auto splat_one = [&](int slot, int count) {
        if (slot >= count) return;
        const float4 c = cand[slot];
        const float px = c.x, py = c.y, pz = c.z, wt = c.w;
        const int x0 = max((int)ceilf(px - r) - T0x, 0), x1 = min((int)floorf(px + r) - T0x, td - 1);
        const int y0 = max((int)ceilf(py - r) - T0y, 0), y1 = min((int)floorf(py + r) - T0y, td - 1);
        const int z0 = max((int)ceilf(pz - r) - T0z, 0), z1 = min((int)floorf(pz + r) - T0z, td - 1);
        for (int z = z0; z <= z1; z++) {
            const float dz = (float)(T0z + z) - pz;
            for (int y = y0; y <= y1; y++) {
                const float dy = (float)(T0y + y) - py;
                for (int x = x0; x <= x1; x++) {
                    const float dx = (float)(T0x + x) - px;
                    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    if (d2 < r2) {
                        const float t = 1.0f - (POW2 ? d2 * inv_r2 : d2 / r2);
                        atomicAdd(&acc[(z * td + y) * td + x], wt * ((t * t) * t));
                    }
                }
            }
        }
    };

    int cnt = 0;
    ChunkIter it;
    bool have = next_bin_chunk(it, sp, cx, cy, cz);
    float4 cur = make_float4(0, 0, 0, 0);
    if (have && it.j0 + lane < it.end) cur = sp.sorted[it.j0 + lane];
    while (have) {
        const int cend = it.end, cj = it.j0 + lane;
        have = next_bin_chunk(it, sp, cx, cy, cz);
        float4 nxt = make_float4(0, 0, 0, 0);
        if (have && it.j0 + lane < it.end) nxt = sp.sorted[it.j0 + lane];
        bool ok = cj < cend;
        {
            const float pw[3] = {cur.x, cur.y, cur.z};
            const int T0[3] = {T0x, T0y, T0z};
#pragma unroll
            for (int a2 = 0; a2 < 3; a2++) {
                const int lo = (int)ceilf(pw[a2] - r) - T0[a2];
                const int hi = (int)floorf(pw[a2] + r) - T0[a2];
                ok = ok && hi >= 0 && lo <= td - 1;
            }
        }
        const uint64_t bal = __ballot(ok);
        if (ok) cand[cnt + (int)__popcll(bal & ((1ull << lane) - 1ull))] = cur;
        cnt += (int)__popcll(bal);
        if (cnt >= 64) {
            __syncthreads();
            splat_one(lane, 64);
            // move the remainder (< 64) to the front
            float4 rest = make_float4(0, 0, 0, 0);
            if (lane < cnt - 64) rest = cand[64 + lane];
            __syncthreads();
            if (lane < cnt - 64) cand[lane] = rest;
            cnt -= 64;
        }
        cur = nxt;
    }
    __syncthreads();
    splat_one(lane, cnt);
    __syncthreads();
    for (int i = lane; i < nv; i += 64) {
        const int x = i % td, y = (i / td) % td, z = i / (td * td);
        brick[((int64_t)(toz + z + a) * br + (toy + y + a)) * br + (tox + x + a)] = acc[i];
    }
}

// ---------------------------------------------------------------------------
// UpdateApron (cuda_gvdb_operators.cuh:73-126) as a neighbour-brick gather:
// apron voxel (x,y,z) of brick k holds the value of world voxel
// mPos + (x,y,z) - 1 in whichever leaf contains it, else boundval.
// ---------------------------------------------------------------------------
// 8^3 bricks with apron 1: the 488 apron voxels of a 10^3 brick in
// ascending brick-linear order, so every lane has one (no interior lanes
// idle) and consecutive lanes write consecutive apron voxels.
struct Apron10Tab {
    uint16_t i[488];
};
constexpr Apron10Tab make_apron10() {
    Apron10Tab t{};
    int n = 0;
    for (int i = 0; i < 1000; i++) {
        const int x = i % 10, y = (i / 10) % 10, z = i / 100;
        if (x == 0 || x == 9 || y == 0 || y == 9 || z == 0 || z == 9) t.i[n++] = (uint16_t)i;
    }
    return t;
}
__constant__ Apron10Tab APRON10 = make_apron10();

// One wave per brick (four per block): the 27 neighbour leaves sit one per
// lane and are fetched with a lane shuffle; each lane's 8 apron voxels are
// loaded before any is stored, so 8 reads per lane are in flight at once.
// Each XCD takes a contiguous run of bricks, so the neighbour faces a brick
// reads were mostly brought into the same L2 by its neighbours.
//
// The x faces are the costly reads: one word in each 40-byte row of the
// neighbour, so a face touches nearly every line of that brick, while a z
// face is one contiguous 320-byte run.  In leaf (key) order the x neighbours
// are a whole y-z slab of cells apart, more bricks than an XCD's L2 holds;
// with ORDER the waves walk the dense level-0 grid with x fastest (empty
// cells exit at once), so a brick's x neighbours were read by the waves just
// before it and its far neighbours are the z ones, whose faces are compact.
//
// PUSHED: a pushing gather (gather_store) has already written every apron
// voxel that has a neighbour leaf; only the boundary word is written, where
// no leaf is (most bricks write nothing).
template <bool ORDER, bool PUSHED>
__global__ __launch_bounds__(256) void k_apron10(ApronParams ap, int64_t nblk) {
    const int lane = threadIdx.x & 63;
    const int64_t blk = gvdbx_xcd_item(blockIdx.x, nblk);
    if (blk < 0) return;
    int64_t k = blk * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if constexpr (ORDER) {
        const CellGrid& g = ap.g0;
        if (k >= g.count()) return;
        const int64_t xy = (int64_t)g.nx * g.ny;
        const int z = (int)(k / xy), y = (int)((k / g.nx) % g.ny), x = (int)(k % g.nx);
        k = ap.idx0[((int64_t)x * g.ny + y) * g.nz + z];
        if (k < 0) return;
    }
    if (k >= ap.nleaf) return;
    int iv[8];
#pragma unroll
    for (int u = 0; u < 8; u++) iv[u] = lane + 64 * u < 488 ? (int)APRON10.i[lane + 64 * u] : -1;
    const GNode nd = ap.leaves[k];
    int32_t leaf = -1;
    if (lane < 27) {
        const int cx = nd.mPos[0] >> 3, cy = nd.mPos[1] >> 3, cz = nd.mPos[2] >> 3;
        const int nx = cx + lane % 3 - 1, ny = cy + (lane / 3) % 3 - 1, nz = cz + lane / 9 - 1;
        const int64_t ncell = (nx >= 0 && ny >= 0 && nz >= 0) ? ap.g0.find(nx, ny, nz) : -1;
        if (ncell >= 0) leaf = ap.idx0[ncell];
        if (leaf >= ap.nleaf) leaf = -1;
    }
    if (PUSHED && !__ballot(lane < 27 && lane != 13 && leaf < 0)) return;   // every neighbour pushed
    // voxels are moved as 4-byte words: float or packed uchar4 channels alike
    const uint32_t* chan = (const uint32_t*)ap.chan;
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int i = iv[u] >= 0 ? iv[u] : 111;   // 111 = (1,1,1): an interior voxel, neighbour 13 (itself)
        const int x = i % 10, y = (i / 10) % 10, z = i / 100;
        const int dx = x == 0 ? -1 : (x == 9 ? 1 : 0);
        const int dy = y == 0 ? -1 : (y == 9 ? 1 : 0);
        const int dz = z == 0 ? -1 : (z == 9 ? 1 : 0);
        const int nl = __shfl(leaf, (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));   // every lane takes part
        if (PUSHED && nl >= 0) iv[u] = -1;   // written by the gather
        // interior voxel of the neighbour holding world voxel mPos + (x,y,z) - 1
        v[u] = ap.bound_bits;
        if (iv[u] >= 0 && nl >= 0) v[u] = chan[(int64_t)nl * 1000 + i - 8 * (dz * 100 + dy * 10 + dx)];
    }
    uint32_t* brick = (uint32_t*)ap.chan + k * 1000;
#pragma unroll
    for (int u = 0; u < 8; u++)
        if (iv[u] >= 0) brick[iv[u]] = v[u];
}

// Per bin cell: the leaf whose cell it is (as k_apron10 finds neighbours:
// idx0 over the level-0 grid, leaves < nleaf), or -1.
__global__ __launch_bounds__(256) void k_bin_leaf(CellGrid gb, CellGrid g0, const int32_t* idx0, int64_t nleaf,
                                                  int32_t* out) {
    const int64_t n = gb.count();
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        gb.cell(c, x, y, z);
        const int64_t f = (x >= 0 && y >= 0 && z >= 0) ? g0.find(x, y, z) : -1;
        const int32_t k = f >= 0 ? idx0[f] : -1;
        out[c] = k < nleaf ? k : -1;
    }
}

// BR > 0: padded brick size known at compile time; BR == 0: runtime size.
template <int BR>
__global__ __launch_bounds__(256) void k_apron(ApronParams ap) {
    const int64_t k = blockIdx.x;
    if (k >= ap.nleaf) return;
    const int br = BR > 0 ? BR : ap.brick_res;
    const int r0 = ap.range0;
    const int a = (br - r0) / 2;
    __shared__ int32_t nbl[27];
    const GNode nd = ap.leaves[k];
    const int cx = nd.mPos[0] >> ap.logr0, cy = nd.mPos[1] >> ap.logr0, cz = nd.mPos[2] >> ap.logr0;
    if (threadIdx.x < 27) {
        int d = threadIdx.x;
        int nx = cx + d % 3 - 1, ny = cy + (d / 3) % 3 - 1, nz = cz + d / 9 - 1;
        int32_t leaf = -1;
        const int64_t ncell = (nx >= 0 && ny >= 0 && nz >= 0) ? ap.g0.find(nx, ny, nz) : -1;
        if (ncell >= 0) leaf = ap.idx0[ncell];
        nbl[d] = (leaf >= 0 && leaf < ap.nleaf) ? leaf : -1;
    }
    __syncthreads();
    const int nv = br * br * br;
    const uint32_t* chan = (const uint32_t*)ap.chan;   // 4-byte words of either channel type
    uint32_t* brick = (uint32_t*)ap.chan + k * (int64_t)nv;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        const int x = i % br, y = (i / br) % br, z = i / (br * br);
        const int dx = x < a ? -1 : (x >= a + r0 ? 1 : 0);
        const int dy = y < a ? -1 : (y >= a + r0 ? 1 : 0);
        const int dz = z < a ? -1 : (z >= a + r0 ? 1 : 0);
        if ((dx | dy | dz) == 0) continue;   // interior voxel
        const int32_t leaf = nbl[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
        uint32_t v = ap.bound_bits;
        if (leaf >= 0)   // interior voxel of the neighbour holding world voxel mPos + (x,y,z) - a
            v = chan[(int64_t)leaf * nv + ((int64_t)(z - dz * r0) * br + (y - dy * r0)) * br + (x - dx * r0)];
        brick[i] = v;
    }
}

// ---------------------------------------------------------------------------
// InsertPoints (cuda_gvdb_particles.cuh:19-36): point -> leaf index and its
// rank among the leaf's points in ascending point order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_insert_points(InsertParams ip) {
    int64_t total = ip.off[ip.gb.count()];
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
        float4 p = ip.sorted[j];
        int i = __float_as_int(p.w);
        int c[3] = {gvdbx_cell(p.x, ip.logr0), gvdbx_cell(p.y, ip.logr0), gvdbx_cell(p.z, ip.logr0)};
        int32_t leaf = -1;
        const int64_t lcell = ip.g0.find(c[0], c[1], c[2]);
        if (lcell >= 0) leaf = ip.idx0[lcell];
        if (leaf >= 0) {
            ip.pnode[i] = leaf;
            ip.poff[i] = (int32_t)(j - ip.off[ip.gb.find(c[0], c[1], c[2])]);   // a leaf's cell is a bin
        }
    }
}

// ---------------------------------------------------------------------------
// ScatterDensity (cuda_gvdb_particles.cuh:525-573) without the surface race:
// each point adds the unclamped cubic distFunc to its home voxel and, with
// expand, to the six face neighbours (which may be apron voxels of the same
// brick).  Each brick's own points are visited in ascending point order, so
// every voxel sums in that order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dist_func(float ax, float ay, float az, float bx, float by, float bz, float r) {
    bx -= ax; by -= ay; bz -= az;
    float c = gvdbx_d2(bx, by, bz) / (r * r);
    float t = 3.0f - c;
    t = c * t;
    t = -3.0f + t;
    t = c * t;
    return 1.0f + t;
}

__global__ __launch_bounds__(128) void k_scatter_density(ScatterParams sp) {
    const int64_t k = blockIdx.x;
    if (k >= sp.nleaf) return;
    const int br = sp.brick_res, a = sp.apron, r0 = sp.res0;
    const GNode nd = sp.leaves[k];
    float* brick = sp.chan + k * (int64_t)br * br * br;
    int c[3] = {nd.mPos[0] / r0, nd.mPos[1] / r0, nd.mPos[2] / r0};
    int beg = 0, end = 0;
    if (const int64_t ci = sp.gb.find(c[0], c[1], c[2]); ci >= 0) {
        beg = sp.off[ci];
        end = sp.off[ci + 1];
    }
    // each thread owns whole z-columns (qx,qy) of the padded brick
    for (int col = threadIdx.x; col < br * br; col += blockDim.x) {
        const int qx = col % br, qy = col / br;
        for (int j = beg; j < end; j++) {
            float4 w = sp.sorted[j];
            float p[3] = {w.x - (float)nd.mPos[0], w.y - (float)nd.mPos[1], w.z - (float)nd.mPos[2]};
            float pi[3] = {(float)(int)p[0], (float)(int)p[1], (float)(int)p[2]};
            if (pi[0] < 0 || pi[1] < 0 || pi[2] < 0 || pi[0] >= r0 || pi[1] >= r0 || pi[2] >= r0) continue;
            int ix = (int)pi[0] + a, iy = (int)pi[1] + a, iz = (int)pi[2] + a;
            float* colp = brick + (int64_t)qy * br + qx;   // voxel z at colp[z*br*br]
            const int64_t zs = (int64_t)br * br;
            if (qx == ix && qy == iy) {
                colp[iz * zs] = colp[iz * zs] + dist_func(p[0], p[1], p[2], pi[0], pi[1], pi[2], sp.radius);
                if (sp.expand) {
                    colp[(iz - 1) * zs] = colp[(iz - 1) * zs] + dist_func(p[0], p[1], p[2], pi[0], pi[1], pi[2] - 1, sp.radius);
                    colp[(iz + 1) * zs] = colp[(iz + 1) * zs] + dist_func(p[0], p[1], p[2], pi[0], pi[1], pi[2] + 1, sp.radius);
                }
            } else if (sp.expand) {
                float d = 0.0f;
                bool hit = true;
                if (qy == iy && qx == ix - 1) d = dist_func(p[0], p[1], p[2], pi[0] - 1, pi[1], pi[2], sp.radius);
                else if (qy == iy && qx == ix + 1) d = dist_func(p[0], p[1], p[2], pi[0] + 1, pi[1], pi[2], sp.radius);
                else if (qx == ix && qy == iy - 1) d = dist_func(p[0], p[1], p[2], pi[0], pi[1] - 1, pi[2], sp.radius);
                else if (qx == ix && qy == iy + 1) d = dist_func(p[0], p[1], p[2], pi[0], pi[1] + 1, pi[2], sp.radius);
                else hit = false;
                if (hit) colp[iz * zs] = colp[iz * zs] + d;
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

inline int res0_of(const CellGrid&, int r0) { return r0; }

}  // namespace

void launch_gather(GatherParams gp, hipStream_t s) {
    int res0 = gp.brick_res - 2 * gp.apron;
    int td = res0 < 8 ? res0 : 8;
    int per = res0 / td;
    int tiles = per * per * per;
    int64_t nblk = gp.nleaf * tiles;
    if (nblk <= 0) return;
    dim3 grid((unsigned)gvdbx_xcd_blocks(nblk));
    const bool halves = td == 8 && gp.sc >= 4;
    if (gp.variant == 0)
        hipLaunchKernelGGL(k_gather, grid, dim3(64), 0, s, gp, res0, tiles, nblk);
    else if (gp.variant == 2)
        hipLaunchKernelGGL(k_gather_pc, grid, dim3(64), 0, s, gp, res0, tiles, nblk);
    else if (gp.coff) {
        if (halves && gp.quarters)
            hipLaunchKernelGGL((k_gather_rl<true, false, true, true>), grid, dim3(64), 0, s, gp, res0, tiles, nblk, GatherSplit{});
        else if (halves)
            hipLaunchKernelGGL((k_gather_rl<true, false, true>), grid, dim3(64), 0, s, gp, res0, tiles, nblk, GatherSplit{});
        else
            hipLaunchKernelGGL((k_gather_rl<false, false, true>), grid, dim3(64), 0, s, gp, res0, tiles, nblk, GatherSplit{});
    } else if (halves)
        hipLaunchKernelGGL((k_gather_rl<true, false, false>), grid, dim3(64), 0, s, gp, res0, tiles, nblk, GatherSplit{});
    else
        hipLaunchKernelGGL((k_gather_rl<false, false, false>), grid, dim3(64), 0, s, gp, res0, tiles, nblk, GatherSplit{});
}

void launch_gather_classes(const float4* sorted, const int32_t* off, CellGrid gb, int res0, float box_radius, int sc,
                           int sc_shift, float cutoff, float4* csorted, int32_t* coff, uint8_t* both, int32_t* ctr,
                           int2* units, int64_t unit_cap, int4* bigs, int64_t big_cap, int32_t* ucnt, hipStream_t s,
                           const Quant16* qz) {
    const Quant16 q = qz ? *qz : Quant16{};
    const int64_t nbin = gb.count();
    if (nbin <= 0) return;
    launch_fill_i32(ctr, 2, 0, s);
    const int logr0 = __builtin_ctz((unsigned)res0);
    hipLaunchKernelGGL(k_gather_classes, dim3((unsigned)std::min<int64_t>(nbin, 1 << 20)), dim3(64), 0, s, sorted, off,
                       gb, logr0, box_radius, sc, sc_shift, cutoff, q, csorted, coff, both, ctr, units, bigs);
    const dim3 gu((unsigned)std::max<int64_t>(1, std::min<int64_t>(unit_cap, 8192)));   // persistent: most bins are small
    hipLaunchKernelGGL(k_class_count, gu, dim3(64), 0, s, sorted, off, logr0, box_radius, sc, sc_shift, cutoff, q, ctr,
                       units, ucnt);
    hipLaunchKernelGGL(k_class_scan, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(big_cap, 2048))), dim3(64),
                       0, s, off, ctr, bigs, ucnt, coff, both);
    hipLaunchKernelGGL(k_class_place, gu, dim3(64), 0, s, sorted, off, logr0, box_radius, sc, sc_shift, cutoff, q, ctr,
                       units, ucnt, csorted);
}

void launch_gather_box(const float4* sorted, const int32_t* total, int64_t nmax, float box_radius, int sc, int sc_shift,
                       float cutoff, uint32_t* box, hipStream_t s) {
    if (nmax > 0)
        hipLaunchKernelGGL(k_gather_box, dim3(grid_for(nmax, 256)), dim3(256), 0, s, sorted, total, nmax, box_radius, sc,
                           sc_shift, cutoff, box);
}

int gather_waves() {
    static int waves = 0;
    if (!waves) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gather_items<true, true>, 64, 0) != hipSuccess)
            cus = 256, per_cu = 8;
        waves = std::max(1, cus) * std::max(1, per_cu);
    }
    return waves;
}

void launch_gather_split(GatherParams gp, int split, int4* items, int64_t item_cap, int32_t* counts,
                         int32_t* slot_tile, int64_t slot_cap, uint32_t* acc, hipStream_t s) {
    const int res0 = gp.brick_res - 2 * gp.apron;
    const int td = res0 < 8 ? res0 : 8;
    const int per = res0 / td;
    const int tiles = per * per * per;
    const int64_t ntiles = gp.nleaf * tiles;
    if (ntiles <= 0) return;
    const bool halves = td == 8 && gp.sc >= 4;
    launch_fill_i32(counts, 2, 0, s);
    const GatherSplit gs{split, items, item_cap, counts, slot_tile, slot_cap, acc};
    const dim3 grid((unsigned)gvdbx_xcd_blocks(ntiles));
    const bool cls = gp.coff != nullptr;
    if (halves && cls && gp.quarters)
        hipLaunchKernelGGL((k_gather_rl<true, true, true, true>), grid, dim3(64), 0, s, gp, res0, tiles, ntiles, gs);
    else if (halves && cls)
        hipLaunchKernelGGL((k_gather_rl<true, true, true>), grid, dim3(64), 0, s, gp, res0, tiles, ntiles, gs);
    else if (halves)
        hipLaunchKernelGGL((k_gather_rl<true, true, false>), grid, dim3(64), 0, s, gp, res0, tiles, ntiles, gs);
    else if (cls)
        hipLaunchKernelGGL((k_gather_rl<false, true, true>), grid, dim3(64), 0, s, gp, res0, tiles, ntiles, gs);
    else
        hipLaunchKernelGGL((k_gather_rl<false, true, false>), grid, dim3(64), 0, s, gp, res0, tiles, ntiles, gs);
    const int waves = gather_waves();
    const dim3 gi((unsigned)std::max<int64_t>(1, std::min<int64_t>(item_cap, waves)));
    if (halves && cls && gp.quarters)
        hipLaunchKernelGGL((k_gather_items<true, true, true>), gi, dim3(64), 0, s, gp, res0, tiles, items, item_cap,
                           counts, acc, slot_cap);
    else if (halves && cls)
        hipLaunchKernelGGL((k_gather_items<true, true>), gi, dim3(64), 0, s, gp, res0, tiles, items, item_cap, counts,
                           acc, slot_cap);
    else if (halves)
        hipLaunchKernelGGL((k_gather_items<true, false>), gi, dim3(64), 0, s, gp, res0, tiles, items, item_cap, counts,
                           acc, slot_cap);
    else if (cls)
        hipLaunchKernelGGL((k_gather_items<false, true>), gi, dim3(64), 0, s, gp, res0, tiles, items, item_cap, counts,
                           acc, slot_cap);
    else
        hipLaunchKernelGGL((k_gather_items<false, false>), gi, dim3(64), 0, s, gp, res0, tiles, items, item_cap, counts,
                           acc, slot_cap);
    const dim3 gf((unsigned)std::max<int64_t>(1, std::min<int64_t>(slot_cap, waves)));
    hipLaunchKernelGGL(k_gather_finish, gf, dim3(64), 0, s, gp, res0, tiles, slot_tile, counts, acc, slot_cap);
}

int splat_halo(float radius) { return (int)std::ceil(radius); }
int splat_width(float radius) { return (int)std::ceil(2.0f * radius); }

int64_t splat_cell_lds(int res0, float radius) {
    const int64_t G = res0 + 2 * splat_halo(radius);
    return G * G * G * 4;
}

void launch_splat_runs(SplatParams sp, int chunk, int32_t* runs, hipStream_t s) {
    const int64_t nc = sp.gb.count();
    if (nc > 0) hipLaunchKernelGGL(k_splat_runs, dim3(grid_for(nc, 256)), dim3(256), 0, s, sp.off, nc, chunk, runs);
}

void launch_splat_run_map(SplatParams sp, const int32_t* run_off, int32_t* run_cell, hipStream_t s) {
    const int64_t nc = sp.gb.count();
    if (nc > 0) hipLaunchKernelGGL(k_splat_run_map, dim3(grid_for(nc, 256)), dim3(256), 0, s, run_off, nc, run_cell);
}

static bool radius_pow2(float radius) {
    int e = 0;
    return std::frexp(radius * radius, &e) == 0.5f;
}

void launch_splat_cell(SplatParams sp, const int32_t* run_cell, const int32_t* run_off, int nruns, int chunk,
                       float* part, hipStream_t s) {
    if (nruns <= 0) return;
    const int res0 = sp.brick_res - 2 * sp.apron;
    const int h = splat_halo(sp.radius), W = splat_width(sp.radius);
    const size_t lds = (size_t)splat_cell_lds(res0, sp.radius);
    const bool pow2 = radius_pow2(sp.radius), small = W <= 4;
    const dim3 grid((unsigned)nruns);
    if (pow2 && small)
        hipLaunchKernelGGL((k_splat_cell<true, true>), grid, dim3(64), lds, s, sp, run_cell, run_off, nruns, res0, h, W,
                           chunk, part);
    else if (pow2)
        hipLaunchKernelGGL((k_splat_cell<true, false>), grid, dim3(64), lds, s, sp, run_cell, run_off, nruns, res0, h,
                           W, chunk, part);
    else if (small)
        hipLaunchKernelGGL((k_splat_cell<false, true>), grid, dim3(64), lds, s, sp, run_cell, run_off, nruns, res0, h,
                           W, chunk, part);
    else
        hipLaunchKernelGGL((k_splat_cell<false, false>), grid, dim3(64), lds, s, sp, run_cell, run_off, nruns, res0, h,
                           W, chunk, part);
}

void launch_splat_cell_backward(SplatBackParams p, CellGrid gb, const int32_t* off, const int32_t* run_cell,
                                const int32_t* run_off, int nruns, int res0, int chunk, hipStream_t s) {
    if (nruns <= 0) return;
    const int h = splat_halo(p.radius), W = splat_width(p.radius);
    const size_t lds = (size_t)splat_cell_lds(res0, p.radius);
    const bool pow2 = radius_pow2(p.radius);
    const dim3 grid((unsigned)nruns);
    if (pow2 && W == 4)
        hipLaunchKernelGGL((k_splat_cell_backward<true, 4>), grid, dim3(64), lds, s, p, gb, off, run_cell, run_off, nruns,
                           res0, h, W, chunk);
    else if (pow2)
        hipLaunchKernelGGL((k_splat_cell_backward<true, 0>), grid, dim3(64), lds, s, p, gb, off, run_cell, run_off, nruns,
                           res0, h, W, chunk);
    else if (W == 4)
        hipLaunchKernelGGL((k_splat_cell_backward<false, 4>), grid, dim3(64), lds, s, p, gb, off, run_cell, run_off,
                           nruns, res0, h, W, chunk);
    else
        hipLaunchKernelGGL((k_splat_cell_backward<false, 0>), grid, dim3(64), lds, s, p, gb, off, run_cell, run_off,
                           nruns, res0, h, W, chunk);
}

void launch_splat_collect(SplatParams sp, const int32_t* run_off, const float* part, hipStream_t s) {
    const int res0 = sp.brick_res - 2 * sp.apron;
    if (sp.nleaf <= 0) return;
    if (res0 == 8)
        hipLaunchKernelGGL(k_splat_collect<8>, dim3((unsigned)sp.nleaf), dim3(256), 0, s, sp, run_off, part, res0,
                           splat_halo(sp.radius));
    else
        hipLaunchKernelGGL(k_splat_collect<0>, dim3((unsigned)sp.nleaf), dim3(256), 0, s, sp, run_off, part, res0,
                           splat_halo(sp.radius));
}

void launch_splat_tile(SplatParams sp, hipStream_t s) {
    const int res0 = sp.brick_res - 2 * sp.apron;
    const int td = res0 < 8 ? res0 : 8;
    const int per = res0 / td;
    const int tiles = per * per * per;
    const int64_t nblk = sp.nleaf * tiles;
    if (nblk <= 0) return;
    const dim3 grid((unsigned)gvdbx_xcd_blocks(nblk));
    if (radius_pow2(sp.radius))
        hipLaunchKernelGGL(k_splat_tile<true>, grid, dim3(64), 0, s, sp, res0, tiles, nblk);
    else
        hipLaunchKernelGGL(k_splat_tile<false>, grid, dim3(64), 0, s, sp, res0, tiles, nblk);
}

void launch_splat_src(const float4* sorted, int64_t n, const int32_t* count, const char* weights, int64_t wstride,
                      float4* out, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_splat_src, dim3(grid_for(n, 256)), dim3(256), 0, s, sorted, n, count, weights, wstride, out);
}

void launch_apron(ApronParams ap, hipStream_t s) {
    if (ap.nleaf <= 0) return;
    const bool ord = ap.order && !ap.g0.sparse();
    const int64_t nb = ord ? (ap.g0.count() + 3) / 4 : (ap.nleaf + 3) / 4;
    const dim3 g10((unsigned)gvdbx_xcd_blocks(nb));
    if (ap.brick_res == 10 && ap.range0 == 8) {
        if (ord && ap.pushed) hipLaunchKernelGGL((k_apron10<true, true>), g10, dim3(256), 0, s, ap, nb);
        else if (ord) hipLaunchKernelGGL((k_apron10<true, false>), g10, dim3(256), 0, s, ap, nb);
        else if (ap.pushed) hipLaunchKernelGGL((k_apron10<false, true>), g10, dim3(256), 0, s, ap, nb);
        else hipLaunchKernelGGL((k_apron10<false, false>), g10, dim3(256), 0, s, ap, nb);
    }
    else if (ap.brick_res == 10)
        hipLaunchKernelGGL(k_apron<10>, dim3((unsigned)ap.nleaf), dim3(256), 0, s, ap);
    else
        hipLaunchKernelGGL(k_apron<0>, dim3((unsigned)ap.nleaf), dim3(256), 0, s, ap);
}

void launch_bin_leaf(CellGrid gb, CellGrid g0, const int32_t* idx0, int64_t nleaf, int32_t* out, hipStream_t s) {
    if (gb.count() > 0)
        hipLaunchKernelGGL(k_bin_leaf, dim3(grid_for(gb.count(), 256)), dim3(256), 0, s, gb, g0, idx0, nleaf, out);
}

void launch_insert_points(InsertParams ip, hipStream_t s) {
    hipLaunchKernelGGL(k_insert_points, dim3(grid_for(ip.n, 256)), dim3(256), 0, s, ip);
}

void launch_scatter_density(ScatterParams sp, hipStream_t s) {
    if (sp.nleaf > 0) hipLaunchKernelGGL(k_scatter_density, dim3((unsigned)sp.nleaf), dim3(128), 0, s, sp);
}
