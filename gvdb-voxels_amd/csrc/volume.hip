// volume.hip -- host runtime and C-ABI of the MI355X point -> sparse-volume
// library.  One gvdbx_volume is one reference nvdb::VolumeGVDB restricted to
// the points path; each exported function names the method it replaces in
// include/gvdbx/gvdbx_c.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "gvdbx/gvdbx_c.h"
#include "gvdbx_kernels.h"

static thread_local char g_err[1024] = "";

void gvdbx_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define FAIL(code, ...)               \
    do {                              \
        gvdbx_set_error(__VA_ARGS__); \
        return (code);                \
    } while (0)

#define TRY(expr)                     \
    do {                              \
        int _s = (expr);              \
        if (_s != GVDBX_OK) return _s; \
    } while (0)

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class T>
    T* as() const { return (T*)p; }
};

constexpr int64_t MAX_DENSE_CELLS = int64_t(1) << 28;

}  // namespace

// Scratch of gvdbx_gather_level_set_noalloc: the packed membership range of
// every binned point (one 32-bit word each).  Device memory that the loader
// reserves on every GPU the library runs on, held until the process ends.
// In host code the name is only a stand-in: the device address comes from
// hipGetSymbolAddress (scratch_on) and is the one that reaches kernels.
__device__ __align__(256) uint32_t scratch[GVDBX_NOALLOC_SCRATCH_BYTES / 4];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (which must be `dev`),
// looked up once per device.
int scratch_on(int dev, uint32_t** out) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        gvdbx_set_error("device %d beyond the %d devices the scratch table holds", dev, SCRATCH_MAX_DEVICES);
        return GVDBX_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            gvdbx_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return GVDBX_ERR_HIP;
        }
        g_scratch_ptr[dev] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return GVDBX_OK;
}

}  // namespace

struct gvdbx_graph;
struct gvdbx_volume;
namespace {
void retire_graphs(gvdbx_volume* v);
}  // namespace

struct gvdbx_volume {
    int device = 0;
    hipStream_t stream = nullptr;
    // configuration (Configure VG.cpp:2380-2434)
    int levs = 0;
    int logdim[GVDBX_MAXLEV] = {};
    int res[GVDBX_MAXLEV] = {};
    int64_t range[GVDBX_MAXLEV] = {};
    bool configured = false;
    // atlas (AtlasCreate AL.cpp:528-560)
    int def_axis[3] = {16, 16, 1};
    int subdim[3] = {0, 0, 0};
    int apron = -1;
    bool has_chan[GVDBX_MAXCHAN] = {};
    int chan_type[GVDBX_MAXCHAN] = {};   // GVDBX_FLOAT or GVDBX_UCHAR4 (4-byte voxels either way)
    int chan_filter[GVDBX_MAXCHAN] = {};
    int chan_border[GVDBX_MAXCHAN] = {};
    DevBuf chan[GVDBX_MAXCHAN];
    int64_t atlas_cap = 0;   // bricks per channel buffer
    // points (SetPoints VG.cpp:4900-4908)
    bool has_pos = false;
    uint64_t pos_gen = 0;   // bumped by SetPoints; bins are reused only within one generation
    PointSrc pos{nullptr, 12};
    const char* wts = nullptr;
    int64_t wstride = 4;
    const char* clr = nullptr;   // uchar4 colour per point (SetPoints clr), or null
    int64_t clr_stride = 4;
    // topology
    bool has_topo = false;
    int depth = 0, top_lev = 1;
    int root_pos[3] = {0, 0, 0};
    float pos_min[3] = {}, pos_max[3] = {}, obj_min[3] = {}, obj_max[3] = {};
    bool bounds_pending = false;   // obj_min/max are in flight to pin + PIN_BOUNDS
    void* pin = nullptr;           // small pinned buffer of its own (bounds readback)
    int64_t cnt[GVDBX_MAXLEV] = {};
    int64_t nbase[GVDBX_MAXLEV] = {};
    CellGrid grid[GVDBX_MAXLEV] = {};
    DevBuf nodes[GVDBX_MAXLEV + 1], clist[GVDBX_MAXLEV + 1], idx[GVDBX_MAXLEV], flag_base[GVDBX_MAXLEV],
        flag_all[GVDBX_MAXLEV];
    DevBuf reach0, nbmask, partials, totals, bbox, misc;
    // incremental updates (ActivateIncreBricksGPU): the next maps, leaf markers, atlas ranks
    DevBuf idx2[GVDBX_MAXLEV], inc_mark, atl_flag, atl_rank;
    bool leaf_inactive = false;   // an incremental update may have deactivated leaves
    bool incre_rebuild = false;   // RequestFullRebuild(false): RebuildTopology updates incrementally
    bool force_full = false;      // RequestFullRebuild(true): the next topology call rebuilds fully
    // grid transform of SetTransform (VG.cpp:5770-5795), in VBX file order:
    // pretranslation, Euler angles, scale, translation
    float xform[12] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0};
    // bumped whenever a device buffer is (re)allocated: a captured frame
    // graph holds the addresses it saw
    uint64_t buf_gen = 0;
    bool bins_force = false;      // ensure_bins sorts even when the bins look current (graph capture)
    // the bins RebuildTopology sorted may serve the next point-reading call
    // once (the reference pipeline RebuildTopology -> InsertPointsSubcell reads
    // the same points twice); every other call reads the points again, since
    // the caller may have rewritten the buffer in place
    bool bins_once = false;
    std::vector<gvdbx_graph*> graphs;   // live frame graphs of this volume (retired by gvdbx_destroy)
    int topo_mode = GVDBX_TOPO_AUTO, topo_mode_used = 0;
    int gather_variant = 1;
    int gather_split = GATHER_SPLIT_DEFAULT;
    int gather_classes = 1;
    int xcd_remap = 0;
    int splat_chunk = SPLAT_CHUNK_DEFAULT;
    int splat_variant = 1;
    int sort_tiles = 0;   // point binning sort: 0 auto (wide from 2^23 valid points), 1 narrow, 2 wide
    int bin_records = 1;  // 1: binning sort moves point records; 0: sorts (cell, index) pairs, then gathers
    int apron_order = 1;  // UpdateApron: 1 = bricks in cell order with x fastest (dense grids), 0 = leaf order
    int apron_push = 1;   // level-set / density gathers on 8^3 bricks also fill the neighbours' aprons
    int gather_quarters = 1;  // 1: the walk prunes z per 2-voxel quarter (8-voxel tiles with reach classes)
    int clr_chan = -1;        // SetColorChannel: the uchar4 channel colouring Render, -1 none
    // per channel: the apron voxels that have a neighbour leaf hold its current
    // values (a pushing gather was the last write, on topology apron_push_gen);
    // chan_raw: the caller holds the device pointer and may write any voxel
    bool apron_pushed[GVDBX_MAXCHAN] = {};
    uint64_t apron_push_gen[GVDBX_MAXCHAN] = {};
    bool chan_raw[GVDBX_MAXCHAN] = {};
    // atlas map
    bool has_atlas = false;
    DevBuf amap;
    int64_t amap_cnt = 0;
    // binning
    DevBuf bin_off, sorted, sorted2, bkey, bval, bkey2, bval2, rs_hist, rs_off, rs_part;
    DevBuf splat_src;   // (x, y, z, weight) in bin order, made per SplatCubic call
    DevBuf gs_items, gs_counts, gs_slots, gs_acc;   // Gather* work items of dense tiles
    DevBuf gbox;                                    // Gather*: packed ranges per bin entry
    DevBuf gcls_pts, gcoff, gcboth;                 // Gather*: bin entries by reach class, class starts, both-face axes
    DevBuf gcctr, gcunit, gcbig, gcucnt;            // Gather*: units of large bins (counters, units, bins, counts)
    DevBuf bin_leaf;                                // Gather* with apron push: leaf of each bin cell
    DevBuf sp_runs, sp_runoff, sp_runcell, sp_part, sp_scan;   // SplatCubic cell push: runs and their grids
    CellGrid gbin = {};
    // sparse topology: per level the sorted cell keys and their hash table,
    // the same for the bin grid, and construction scratch
    DevBuf sg_keys[GVDBX_MAXLEV], sg_ht[GVDBX_MAXLEV], sg_bkeys, sg_bht;
    DevBuf sg_k1, sg_k2, sg_v1, sg_v2, sg_flag, sg_off, sg_part;
    CellGrid gb_sparse = {};   // bin grid of a sparse topology
    bool bins_valid = false;
    int64_t bins_n = -1;
    int64_t pts_valid = -1;   // valid points of the last RebuildTopology (-1: unknown)
    float bins_trans[3] = {};
    const char* bins_base = nullptr;
    int64_t bins_stride = 0;
    uint64_t bins_gen = 0;
    // subcell state (InsertPointsSubcell)
    int sc_size = 0;
    float sc_radius = 0.0f;
    int64_t sc_len = 0;
    bool sc_valid = false;
    bool sc_fp16 = false;   // lists of InsertPointsSubcell_FP16 (positions as 16-bit fractions of sc_q's box)
    Quant16 sc_q = {};
    // InsertPoints output
    DevBuf pnode, poff;
    int64_t npnode = 0;
    // recorded emission-absorption segments (gvdbx_raycast_ea with record)
    DevBuf vis_key, vis_rec, vis_idx, vis_tkey, vis_tidx, vis_hist, vis_off, vis_part, vis_seg, vis_cnt;
    DevBuf vis_lkey, vis_ctag, vis_cbase, vis_csize, vis_coff;   // checkpoint chunks
    int64_t vis_nchunk = 0, vis_ccap = 0;
    int64_t vis_nwaves = 0;   // waves of the recorded launch (their record totals follow vis_cnt's counters)
    int64_t vis_n = 0, vis_cap = 0;
    int vis_leaf_bits = 0;
    bool vis_valid = false, vis_sorted = false, vis_in_tmp = false;
    int vis_w = 0, vis_h = 0, vis_chan = -1, vis_ncam = 0;
    uint64_t vis_topo_gen = 0;
    uint64_t topo_gen = 0;   // bumped by every RebuildTopology
    // batched clouds (gvdbx_set_points_batched): cloud c owns points
    // [pt_off[c], pt_off[c+1]), level-0 cells shifted by cell_off[3c..3c+2]
    int ncloud = 0;
    std::vector<int64_t> h_ptoff;
    std::vector<int32_t> h_celloff, h_boxes;
    std::vector<float> h_cbounds;   // 6 per cloud, cloud-local leaf bounds
    DevBuf c_ptoff, c_celloff, c_boxes, c_tiles, c_cellcloud, c_leafcloud, c_bbox, c_bounds, c_cams;
    // pinned host staging
    void* host = nullptr;
    size_t host_cap = 0;
};

namespace {

int set_dev(gvdbx_volume* v) {
    hipError_t e = hipSetDevice(v->device);
    if (e != hipSuccess) FAIL(GVDBX_ERR_HIP, "hipSetDevice(%d): %s", v->device, hipGetErrorString(e));
    return GVDBX_OK;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) FAIL(GVDBX_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return GVDBX_OK;
}

int sync(gvdbx_volume* v) {
    hipError_t e = hipStreamSynchronize(v->stream);
    if (e != hipSuccess) FAIL(GVDBX_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
    return GVDBX_OK;
}

int ensure(gvdbx_volume* v, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return GVDBX_OK;
    TRY(sync(v));
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 4 + 256;
    v->buf_gen++;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        FAIL(GVDBX_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    return GVDBX_OK;
}

// ensure() that keeps the first `keep` bytes when the buffer grows
int ensure_keep(gvdbx_volume* v, DevBuf& b, size_t bytes, size_t keep) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return GVDBX_OK;
    TRY(sync(v));
    size_t want = bytes + bytes / 4 + 256;
    void* np = nullptr;
    v->buf_gen++;
    hipError_t e = hipMalloc(&np, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        FAIL(GVDBX_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
    }
    if (b.p && keep > 0) e = hipMemcpyAsync(np, b.p, std::min(keep, b.cap), hipMemcpyDeviceToDevice, v->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v->stream);
    if (e != hipSuccess) {
        (void)hipFree(np);
        FAIL(GVDBX_ERR_HIP, "pool growth copy failed: %s", hipGetErrorString(e));
    }
    if (b.p) (void)hipFree(b.p);
    b.p = np;
    b.cap = want;
    return GVDBX_OK;
}

// Layout of gvdbx_volume::pin: the six leaf-bound integers read back by
// FinishTopology (never read while the copy may be in flight: the bounds
// resolution synchronises first).
constexpr size_t PIN_BOUNDS = 0, PIN_BYTES = 64;

int ensure_pin(gvdbx_volume* v) {
    if (v->pin) return GVDBX_OK;
    hipError_t e = hipHostMalloc(&v->pin, PIN_BYTES, hipHostMallocDefault);
    if (e != hipSuccess) {
        v->pin = nullptr;
        FAIL(GVDBX_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", PIN_BYTES, hipGetErrorString(e));
    }
    return GVDBX_OK;
}

// Bounds computed by FinishTopology come back asynchronously; the first
// reader waits for them.
int resolve_bounds(gvdbx_volume* v) {
    if (!v->bounds_pending) return GVDBX_OK;
    TRY(sync(v));
    int32_t out[6];
    memcpy(out, (const char*)v->pin + PIN_BOUNDS, sizeof(out));
    for (int a = 0; a < 3; a++) {
        v->obj_min[a] = (float)out[a];
        v->obj_max[a] = (float)out[3 + a];
    }
    v->bounds_pending = false;
    return GVDBX_OK;
}

int ensure_host(gvdbx_volume* v, size_t bytes) {
    if (v->host_cap >= bytes) return GVDBX_OK;
    TRY(sync(v));
    if (v->host) (void)hipHostFree(v->host);
    v->host = nullptr;
    v->host_cap = 0;
    size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipHostMalloc(&v->host, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        v->host = nullptr;
        FAIL(GVDBX_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    v->host_cap = want;
    return GVDBX_OK;
}

void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Every write of a channel other than a pushing gather leaves its aprons stale.
void chan_written(gvdbx_volume* v, int chan) {
    if (chan >= 0 && chan < GVDBX_MAXCHAN) v->apron_pushed[chan] = false;
}

int brick_res(const gvdbx_volume* v) { return v->res[0] + 2 * std::max(v->apron, 0); }
int64_t brick_vox(const gvdbx_volume* v) {
    int64_t br = brick_res(v);
    return br * br * br;
}

CloudMap cloud_map(const gvdbx_volume* v) {
    CloudMap cm{};
    cm.ncloud = v->ncloud;
    cm.pt_off = v->ncloud ? v->c_ptoff.as<int64_t>() : nullptr;
    cm.cell_off = v->ncloud ? v->c_celloff.as<int32_t>() : nullptr;
    return cm;
}

bool any_chan(const gvdbx_volume* v) {
    for (int c = 0; c < GVDBX_MAXCHAN; c++)
        if (v->has_chan[c]) return true;
    return false;
}

// Grow every channel to hold `bricks` bricks, preserving contents (AtlasResize AL.cpp:622-650).
int grow_atlas(gvdbx_volume* v, int64_t bricks) {
    if (bricks <= v->atlas_cap) return GVDBX_OK;
    // physical capacity grows by at least a quarter, so a sequence of
    // slightly larger topologies reallocates (and copies) rarely; the
    // logical atlas (subdim, node mValue) is what UpdateAtlas computes
    if (v->atlas_cap > 0) bricks = std::max(bricks, v->atlas_cap + v->atlas_cap / 4);
    int64_t bv = brick_vox(v);
    TRY(sync(v));
    v->buf_gen++;
    for (int c = 0; c < GVDBX_MAXCHAN; c++) {
        if (!v->has_chan[c]) continue;
        chan_written(v, c);
        void* np = nullptr;
        size_t bytes = (size_t)(bricks * bv) * sizeof(float);
        hipError_t e = hipMalloc(&np, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            FAIL(GVDBX_ERR_NOMEM, "atlas growth to %lld bricks failed: %s", (long long)bricks, hipGetErrorString(e));
        }
        launch_fill_f32((float*)np, bricks * bv, 0.0f, v->stream);
        if (v->chan[c].p && v->atlas_cap > 0)
            GVDBX_CHECK(hipMemcpyAsync(np, v->chan[c].p, (size_t)(v->atlas_cap * bv) * sizeof(float),
                                       hipMemcpyDeviceToDevice, v->stream));
        TRY(sync(v));
        free_buf(v->chan[c]);
        v->chan[c].p = np;
        v->chan[c].cap = bytes;
    }
    v->atlas_cap = bricks;
    return GVDBX_OK;
}

int check_vol(gvdbx_volume* v) {
    if (!v) FAIL(GVDBX_ERR_ARG, "null volume handle");
    return set_dev(v);
}

int check_chan(gvdbx_volume* v, int chan) {
    if (chan < 0 || chan >= GVDBX_MAXCHAN || !v->has_chan[chan]) FAIL(GVDBX_ERR_ARG, "channel %d does not exist", chan);
    return GVDBX_OK;
}

int check_chan_type(gvdbx_volume* v, int chan, int dtype, const char* what) {
    TRY(check_chan(v, chan));
    if (v->chan_type[chan] != dtype)
        FAIL(GVDBX_ERR_ARG, "%s: channel %d is a %s channel, this call needs a %s one", what, chan,
             v->chan_type[chan] == GVDBX_UCHAR4 ? "uchar4" : "float", dtype == GVDBX_UCHAR4 ? "uchar4" : "float");
    return GVDBX_OK;
}

// float -> uchar as CUDA's saturating conversion (truncation, clamp, NaN -> 0)
uint32_t sat_u8(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 255.0f) return 255u;
    return (uint32_t)x;
}

// Host restatement of DetermineDepth (VG.cpp:821-837) over GetCoveringNode (VG.cpp:2915-2931).
int determine_depth(const gvdbx_volume* v, const float pmin[3], const float pmax[3], int root[3]) {
    int ipmin[3] = {(int)pmin[0], (int)pmin[1], (int)pmin[2]};
    int ipmax[3] = {(int)pmax[0], (int)pmax[1], (int)pmax[2]};
    for (int lev = 0; lev < v->levs; lev++) {
        int64_t rg = v->range[lev];
        int64_t a0[3], a1[3];
        for (int a = 0; a < 3; a++) {
            a0[a] = ((int64_t)ipmin[a] / rg) * rg;
            if (ipmin[a] < a0[a]) a0[a] -= rg;
            a1[a] = ((int64_t)ipmax[a] / rg) * rg;
            if (ipmax[a] < a1[a]) a1[a] -= rg;
        }
        if (a0[0] == a1[0] && a0[1] == a1[1] && a0[2] == a1[2]) {
            for (int a = 0; a < 3; a++) root[a] = (int)a0[a];
            return lev;
        }
    }
    return -1;
}

// Bin the current points by home brick cell: a stable radix sort of
// (cell, point index) over the leaf grid grown by one cell, so every bin
// holds its points in ascending index order.  Bins made by RebuildTopology
// or an earlier call are reused while SetPoints has not been called again and
// the point count and translation match.
// With rp, points binned by this call also mark their extra-brick reach in
// reach0 (*reached set); bins that are reused do not.
int ensure_bins(gvdbx_volume* v, int64_t n, const float t[3], bool stable, const ReachParams* rp = nullptr,
                bool* reached = nullptr) {
    if (reached) *reached = false;
    (void)stable;   // sorted bins are always in ascending point order
    if (!v->has_pos) FAIL(GVDBX_ERR_STATE, "SetPoints has not been called");
    const CellGrid& g0 = v->grid[0];
    // the leaf grid grown by one cell (a sparse topology lists its bin cells)
    const CellGrid gb = g0.sparse() ? v->gb_sparse : CellGrid{g0.ox - 1, g0.oy - 1, g0.oz - 1, g0.nx + 2, g0.ny + 2, g0.nz + 2};
    int64_t nb = gb.count();
    if (nb >= (int64_t)UINT32_MAX) FAIL(GVDBX_ERR_RANGE, "bin grid exceeds 32-bit cell keys");
    bool same = v->bins_valid && v->bins_gen == v->pos_gen && v->bins_n == n && v->bins_base == v->pos.base &&
                v->bins_stride == v->pos.stride && v->bins_trans[0] == t[0] && v->bins_trans[1] == t[1] &&
                v->bins_trans[2] == t[2] && !memcmp(&v->gbin, &gb, sizeof(gb));
    const bool reuse = same && v->bins_once && !v->bins_force;
    v->bins_once = false;
    if (reuse) return GVDBX_OK;
    hipStream_t s = v->stream;
    const int64_t m = std::max<int64_t>(n, 1);
    const int64_t nh = radix_hist_entries(m);
    TRY(ensure(v, v->bin_off, (size_t)(nb + 1) * 4));
    if (v->ncloud == 0 && v->bin_records && !gb.sparse()) {
        // one cloud: the sort moves the point records (no random point reads;
        // each pass finds a record's bin again, which a sparse grid makes a
        // hash probe, so sparse grids sort (bin, index) pairs instead)
        TRY(ensure(v, v->bkey, (size_t)m * 4));
        TRY(ensure(v, v->rs_hist, (size_t)nh * 4));
        TRY(ensure(v, v->rs_off, (size_t)(nh + 1) * 4));
        TRY(ensure(v, v->rs_part, (size_t)(scan_tiles(nh + 1) + 1) * 4));
        TRY(ensure(v, v->sorted, (size_t)m * 16));
        TRY(ensure(v, v->sorted2, (size_t)m * 16));
        const int which = bin_sort_records(v->pos, n, t[0], t[1], t[2], gb, v->logdim[0], v->sorted.as<float4>(),
                                           v->sorted2.as<float4>(), v->bkey.as<uint32_t>(), v->rs_hist.as<int32_t>(),
                                           v->rs_off.as<int32_t>(), v->rs_part.as<uint32_t>(), rp,
                                           v->reach0.as<uint32_t>(), s);
        if (which == 1) std::swap(v->sorted, v->sorted2);
        if (reached) *reached = rp != nullptr && n > 0;
        launch_bin_offsets(v->bkey.as<uint32_t>(), n, nb, v->bin_off.as<int32_t>(), s);
        TRY(check_launch("bin points"));
        v->gbin = gb;
        v->bins_valid = true;
        v->bins_n = n;
        v->bins_gen = v->pos_gen;
        v->bins_base = v->pos.base;
        v->bins_stride = v->pos.stride;
        memcpy(v->bins_trans, t, sizeof(v->bins_trans));
        return GVDBX_OK;
    }
    TRY(ensure(v, v->bkey, (size_t)m * 4));
    TRY(ensure(v, v->bval, (size_t)m * 4));
    TRY(ensure(v, v->bkey2, (size_t)m * 4));
    TRY(ensure(v, v->bval2, (size_t)m * 4));
    TRY(ensure(v, v->rs_hist, (size_t)nh * 4));
    TRY(ensure(v, v->rs_off, (size_t)(nh + 1) * 4));
    TRY(ensure(v, v->rs_part, (size_t)(scan_tiles(nh + 1) + 1) * 4));
    TRY(ensure(v, v->sorted, (size_t)m * 16));
    int end_bit = 1;
    while (((int64_t)1 << end_bit) <= nb) end_bit++;
    launch_bin_keys(v->pos, n, t[0], t[1], t[2], gb, v->logdim[0], cloud_map(v), v->bkey.as<uint32_t>(), nullptr, s);
    int flip = radix_sort_pairs32(v->bkey.as<uint32_t>(), v->bval.as<uint32_t>(), v->bkey2.as<uint32_t>(),
                                  v->bval2.as<uint32_t>(), n, 0, end_bit, v->rs_hist.as<int32_t>(),
                                  v->rs_off.as<int32_t>(), v->rs_part.as<uint32_t>(), s, true,
                                  v->sort_tiles == 2 ||
                                      (v->sort_tiles == 0 && (v->pts_valid >= 0 ? v->pts_valid : n) >= ((int64_t)1 << 23)));
    // values: point indices; wide tiles for 2^23 valid points or more (or as sort_tiles says)
    const uint32_t* skey = flip ? v->bkey2.as<uint32_t>() : v->bkey.as<uint32_t>();
    const uint32_t* sval = flip ? v->bval2.as<uint32_t>() : v->bval.as<uint32_t>();
    launch_bin_gather(v->pos, n, t[0], t[1], t[2], skey, sval, (uint32_t)nb, v->sorted.as<float4>(), rp,
                      v->reach0.as<uint32_t>(), s);
    if (reached) *reached = rp != nullptr;
    launch_bin_offsets(skey, n, nb, v->bin_off.as<int32_t>(), s);
    TRY(check_launch("bin points"));
    v->gbin = gb;
    v->bins_valid = true;
    v->bins_n = n;
    v->bins_gen = v->pos_gen;
    v->bins_base = v->pos.base;
    v->bins_stride = v->pos.stride;
    memcpy(v->bins_trans, t, sizeof(v->bins_trans));
    return GVDBX_OK;
}

// InsertPointsSubcell list length over the current bins.
int subcell_len(gvdbx_volume* v, int sc, float radius, int64_t* len_out) {
    hipStream_t s = v->stream;
    const CellGrid& g0 = v->grid[0];
    TRY(ensure(v, v->nbmask, (size_t)std::max<int64_t>(g0.count(), 1) * 4));
    TRY(ensure(v, v->misc, 256));
    TRY(ensure_host(v, 4096));
    unsigned long long* len = (unsigned long long*)((char*)v->misc.p + 64);
    launch_leaf_nbmask(v->idx[0].as<int32_t>(), g0, v->nbmask.as<uint32_t>(), len, s);   // also len = 0
    BinParams bp{};
    bp.gb = v->gbin;
    bp.range0 = v->res[0];
    bp.logr0 = v->logdim[0];
    bp.sc = sc;
    bp.sc_shift = __builtin_ctz((unsigned)sc);
    bp.radius = radius;
    bp.g0 = g0;
    bp.nbmask = v->nbmask.as<uint32_t>();
    launch_subcell_len(v->sorted.as<float4>(), v->bin_off.as<int32_t>() + v->gbin.count(), v->bins_n, bp, len, s);
    TRY(check_launch("subcell length"));
    GVDBX_CHECK(hipMemcpyAsync(v->host, len, 8, hipMemcpyDeviceToHost, s));
    TRY(sync(v));
    *len_out = (int64_t)(*(unsigned long long*)v->host);
    return GVDBX_OK;
}

// GetBoundingBox (VG.cpp:5452-5463) of one accumulator, DetermineDepth and
// the level-0 cell box the topology may occupy (the root node clipped to the
// points' box grown by the extra-brick reach).  Shared by one volume and by
// each cloud of a batch.
struct CloudFrame {
    float mn[3], mx[3], pmin[3], pmax[3];
    int depth;
    int root[3];
    int lo_v[3], hi_v[3];   // voxel box, inclusive
};

int frame_of(const gvdbx_volume* v, const BBoxAccum& bb, float radius, const char* who, CloudFrame& f) {
    if (bb.nvalid == 0) FAIL(GVDBX_ERR_NOPOINTS, "RebuildTopology: no valid point%s", who);
    for (int a = 0; a < 3; a++) {
        memcpy(&f.mn[a], &bb.mn[a], 4);
        memcpy(&f.mx[a], &bb.mx[a], 4);
        f.pmin[a] = f.mn[a] - 1;
        f.pmax[a] = f.mx[a] + 2;
        f.pmin[a] = 0 > f.pmin[a] ? 0 : f.pmin[a];
        f.pmax[a] = 0 > f.pmax[a] ? 0 : f.pmax[a];
    }
    f.depth = determine_depth(v, f.pmin, f.pmax, f.root);
    if (f.depth == 0) FAIL(GVDBX_ERR_DEPTH, "RebuildTopology: point extents smaller than one brick%s", who);
    if (f.depth < 0 || f.depth >= v->levs)
        FAIL(GVDBX_ERR_DEPTH, "RebuildTopology: point extents exceed the configured levels%s", who);
    for (int a = 0; a < 3; a++)
        if ((int64_t)f.pmax[a] / v->range[0] > 65535)
            FAIL(GVDBX_ERR_RANGE, "RebuildTopology: extents exceed 16-bit brick keys%s", who);
    const bool extra = radius >= 1.0f;   // VG.cpp:1207-1208
    const int64_t rtop = v->range[f.depth];
    for (int a = 0; a < 3; a++) {
        int fmn = (int)f.mn[a], fmx = (int)f.mx[a];
        if (extra) {
            f.lo_v[a] = (int)std::max<int64_t>(f.root[a], (int64_t)fmn - (int64_t)ceilf(radius) - 1);
            f.hi_v[a] = (int)std::min<int64_t>(f.root[a] + rtop - 1, (int64_t)fmx + (int64_t)ceilf(2.0f * radius) + 1);
        } else {
            f.lo_v[a] = std::max(f.root[a], fmn);
            f.hi_v[a] = (int)std::min<int64_t>(f.root[a] + rtop - 1, fmx);
        }
    }
    return GVDBX_OK;
}

// ---- sparse cell grids (sparse.hip) ----
// Sorts the n keys in sg_k1 (packed keys of `bits` bits; `none` = 2^bits
// sorts last) with the LSD radix sort, flags the head of every run, and
// stores the distinct keys, in order, in `out` (their ranks are the prefix
// sum of the flags).
int sg_sort_unique(gvdbx_volume* v, int64_t n, int bits, DevBuf& out, int64_t* count) {
    hipStream_t s = v->stream;
    *count = 0;
    if (n <= 0) return ensure(v, out, 8);
    const uint64_t none = 1ull << bits;
    const int64_t nh = radix_hist_entries(n);
    const int64_t st = scan_tiles(n + 1);
    TRY(ensure(v, v->sg_k2, (size_t)n * 8));
    TRY(ensure(v, v->sg_v1, (size_t)n * 4));
    TRY(ensure(v, v->sg_v2, (size_t)n * 4));
    TRY(ensure(v, v->rs_hist, (size_t)nh * 4));
    TRY(ensure(v, v->rs_off, (size_t)(nh + 1) * 4));
    TRY(ensure(v, v->rs_part, (size_t)(scan_tiles(nh + 1) + 1) * 4));
    TRY(ensure(v, v->sg_flag, (size_t)(n + 1) * 4));
    TRY(ensure(v, v->sg_off, (size_t)(n + 1) * 4));
    TRY(ensure(v, v->sg_part, (size_t)(st + 1) * 4));
    TRY(ensure_host(v, 4096));
    const int flip = radix_sort_pairs(v->sg_k1.as<uint64_t>(), v->sg_v1.as<uint32_t>(), v->sg_k2.as<uint64_t>(),
                                      v->sg_v2.as<uint32_t>(), n, 0, bits + 1, v->rs_hist.as<int32_t>(),
                                      v->rs_off.as<int32_t>(), v->rs_part.as<uint32_t>(), s, true);
    const uint64_t* sk = flip ? v->sg_k2.as<uint64_t>() : v->sg_k1.as<uint64_t>();
    int32_t* flag = v->sg_flag.as<int32_t>();
    int32_t* off = v->sg_off.as<int32_t>();
    uint32_t* part = v->sg_part.as<uint32_t>();
    launch_sp_heads(sk, n, none, flag, s);
    launch_count_scan_reduce(flag, n, part, part + st, s);
    launch_count_scan_apply(flag, n, part, off, s);
    TRY(check_launch("sparse cell keys"));
    GVDBX_CHECK(hipMemcpyAsync(v->host, part + st, 4, hipMemcpyDeviceToHost, s));
    TRY(sync(v));
    uint32_t m;
    memcpy(&m, v->host, 4);
    *count = m;
    TRY(ensure(v, out, (size_t)std::max<int64_t>(m, 1) * 8));
    launch_sp_compact(sk, flag, off, n, out.as<uint64_t>(), s);
    return check_launch("sparse cell keys");
}

// Hash table of the `count` keys in `keys` (rank i for key i) into ht; sets
// the grid's sparse fields.
int sg_table(gvdbx_volume* v, DevBuf& keys, int64_t count, DevBuf& ht, CellGrid& g) {
    uint32_t cap = 64;
    while ((int64_t)cap < 2 * count) cap <<= 1;
    TRY(ensure(v, ht, (size_t)cap * 16));
    launch_fill_u64(ht.as<uint64_t>(), (int64_t)cap * 2, ~0ull, v->stream);
    launch_sp_hash(keys.as<uint64_t>(), count, ht.as<uint4>(), cap - 1, v->stream);
    TRY(check_launch("sparse cell table"));
    g.ht = ht.as<uint4>();
    g.keys = keys.as<uint64_t>();
    g.ncell = count;
    g.hmask = cap - 1;
    return GVDBX_OK;
}

// Sparse grids for a topology whose level-0 box is too large for dense
// per-cell arrays.  The boxes stay those of the dense grids; only the listed
// cells get per-cell storage.  Level 0 lists the points' home cells dilated
// by the extra-brick reach (a superset of every leaf the topology can make:
// a point's box of gvdbCalcExtraBrickId, cuda_gvdb_particles.cuh:493-494,
// spans cells home - ceil(ceil(r)/r0) .. home + floor((r0 - 1 + ceil(2r))/r0));
// each level above lists the parents of the one below; the bin grid lists
// the level-0 cells dilated by one more cell (every bin a leaf's
// neighbourhood reads).
int build_sparse_grids(gvdbx_volume* v, int64_t n, const float orig[3], float radius, bool extra, int depth,
                       CellGrid* grid, CellGrid& gb) {
    hipStream_t s = v->stream;
    const int r0 = v->res[0];
    int bits[GVDBX_MAXLEV];
    for (int l = 0; l < depth; l++) {
        CellGrid& g = grid[l];
        g.by = gvdbx_bits_for(g.ny);
        g.bz = gvdbx_bits_for(g.nz);
        bits[l] = gvdbx_bits_for(g.nx) + g.by + g.bz;
    }
    gb = CellGrid{grid[0].ox - 1, grid[0].oy - 1, grid[0].oz - 1, grid[0].nx + 2, grid[0].ny + 2, grid[0].nz + 2};
    gb.by = gvdbx_bits_for(gb.ny);
    gb.bz = gvdbx_bits_for(gb.nz);
    const int bitsb = gvdbx_bits_for(gb.nx) + gb.by + gb.bz;
    const int64_t lim = ((int64_t)1 << 31) - 1;
    // home cells of the points
    TRY(ensure(v, v->sg_k1, (size_t)std::max<int64_t>(n, 1) * 8));
    launch_sp_point_keys(v->pos, n, orig[0], orig[1], orig[2], grid[0], v->logdim[0], 1ull << bits[0],
                         v->sg_k1.as<uint64_t>(), s);
    int64_t nhome = 0;
    TRY(sg_sort_unique(v, n, bits[0], v->sg_keys[0], &nhome));
    // level 0: dilated by the extra-brick reach
    int64_t nc[GVDBX_MAXLEV] = {};
    nc[0] = nhome;
    if (extra) {
        const int cr = (int)ceilf(radius), c2r = (int)ceilf(2.0f * radius);
        const int mlo = (cr + r0 - 1) / r0, mhi = (r0 - 1 + c2r) / r0;
        const int w = mlo + mhi + 1;
        const int64_t nd = nhome * w * w * w;
        if (nd > lim) FAIL(GVDBX_ERR_RANGE, "RebuildTopology: %lld candidate cells exceed the sparse-grid limit", (long long)nd);
        TRY(ensure(v, v->sg_k1, (size_t)std::max<int64_t>(nd, 1) * 8));
        launch_sp_dilate(v->sg_keys[0].as<uint64_t>(), nhome, grid[0], -mlo, mhi, grid[0], 1ull << bits[0],
                         v->sg_k1.as<uint64_t>(), s);
        TRY(sg_sort_unique(v, nd, bits[0], v->sg_keys[0], &nc[0]));
    }
    // bin grid: one more cell around
    {
        const int64_t nd = nc[0] * 27;
        if (nd > lim) FAIL(GVDBX_ERR_RANGE, "RebuildTopology: %lld bin cells exceed the sparse-grid limit", (long long)nd);
        TRY(ensure(v, v->sg_k1, (size_t)std::max<int64_t>(nd, 1) * 8));
        launch_sp_dilate(v->sg_keys[0].as<uint64_t>(), nc[0], grid[0], -1, 1, gb, 1ull << bitsb,
                         v->sg_k1.as<uint64_t>(), s);
    }
    int64_t nbin = 0;
    TRY(sg_sort_unique(v, nc[0] * 27, bitsb, v->sg_bkeys, &nbin));
    // levels above: parents
    for (int l = 1; l < depth; l++) {
        TRY(ensure(v, v->sg_k1, (size_t)std::max<int64_t>(nc[l - 1], 1) * 8));
        launch_sp_parents(v->sg_keys[l - 1].as<uint64_t>(), nc[l - 1], grid[l - 1], v->logdim[l], grid[l],
                          1ull << bits[l], v->sg_k1.as<uint64_t>(), s);
        TRY(sg_sort_unique(v, nc[l - 1], bits[l], v->sg_keys[l], &nc[l]));
    }
    for (int l = 0; l < depth; l++) TRY(sg_table(v, v->sg_keys[l], nc[l], v->sg_ht[l], grid[l]));
    return sg_table(v, v->sg_bkeys, nbin, v->sg_bht, gb);
}

int rebuild_topology(gvdbx_volume* v, int64_t n, float radius, const float orig[3]) {
    if (!v->configured) FAIL(GVDBX_ERR_STATE, "Configure has not been called");
    if (!v->has_pos) FAIL(GVDBX_ERR_STATE, "SetPoints has not been called");
    if (n < 0) FAIL(GVDBX_ERR_ARG, "negative point count");
    if (!(radius >= 0.0f) || !isfinite(radius)) FAIL(GVDBX_ERR_ARG, "radius must be finite and >= 0");
    hipStream_t s = v->stream;
    const int r0 = v->res[0];
    const bool extra = radius >= 1.0f;   // VG.cpp:1207-1208
    const bool mask_ok = 2.0f * radius <= (float)r0 && radius + 1.0f <= (float)r0;
    int depth;
    int root[3];
    float pmin[3], pmax[3];
    int lo_v[3], hi_v[3];
    if (v->ncloud == 0) {
        TRY(ensure(v, v->bbox, sizeof(BBoxAccum) * (1 + (size_t)bbox_parts(std::max<int64_t>(n, 1)))));
        TRY(ensure_host(v, 4096));
        // the bounding-box pass writes every field of the accumulator; only an
        // empty point set needs the initial (empty) box
        if (n > 0) launch_bbox(v->pos, n, orig[0], orig[1], orig[2], v->bbox.as<BBoxAccum>(), v->bbox.as<BBoxAccum>() + 1, s);
        else launch_bbox_init(v->bbox.as<BBoxAccum>(), s);
        TRY(check_launch("bounding box"));
        GVDBX_CHECK(hipMemcpyAsync(v->host, v->bbox.p, sizeof(BBoxAccum), hipMemcpyDeviceToHost, s));
        TRY(sync(v));
        BBoxAccum bb;
        memcpy(&bb, v->host, sizeof(bb));
        CloudFrame f;
        TRY(frame_of(v, bb, radius, "", f));
        v->pts_valid = (int64_t)bb.nvalid;
        depth = f.depth;
        memcpy(root, f.root, sizeof(root));
        memcpy(pmin, f.pmin, sizeof(pmin));
        memcpy(pmax, f.pmax, sizeof(pmax));
        memcpy(lo_v, f.lo_v, sizeof(lo_v));
        memcpy(hi_v, f.hi_v, sizeof(hi_v));
    } else {
        // ---- batched clouds: per-cloud frames, then one packed index space ----
        if (orig[0] != 0.0f || orig[1] != 0.0f || orig[2] != 0.0f)
            FAIL(GVDBX_ERR_ARG, "RebuildTopology: batched clouds take no origin shift");
        if (extra && !mask_ok) FAIL(GVDBX_ERR_ARG, "RebuildTopology: batched clouds need radius <= half a brick");
        if (n != v->h_ptoff[v->ncloud]) FAIL(GVDBX_ERR_ARG, "RebuildTopology: point count differs from the batch");
        const int nc = v->ncloud;
        std::vector<int32_t> tiles;
        for (int c = 0; c < nc; c++)
            for (int64_t o = 0; o < v->h_ptoff[c + 1] - v->h_ptoff[c]; o += BBOX_CHUNK) {
                tiles.push_back(c);
                tiles.push_back((int32_t)o);
            }
        const int64_t ntiles = (int64_t)tiles.size() / 2;
        TRY(ensure(v, v->c_bbox, (size_t)nc * sizeof(BBoxAccum)));
        TRY(ensure(v, v->c_tiles, std::max<size_t>(tiles.size(), 2) * 4));
        TRY(ensure_host(v, std::max<size_t>(tiles.size() * 4, (size_t)nc * sizeof(BBoxAccum)) + 4096));
        memcpy(v->host, tiles.data(), tiles.size() * 4);
        GVDBX_CHECK(hipMemcpyAsync(v->c_tiles.p, v->host, tiles.size() * 4, hipMemcpyHostToDevice, s));
        launch_bbox_clouds_init(v->c_bbox.as<BBoxAccum>(), nc, s);
        launch_bbox_clouds(v->pos, v->c_tiles.as<int32_t>(), ntiles, v->c_ptoff.as<int64_t>(), v->c_bbox.as<BBoxAccum>(), s);
        TRY(check_launch("cloud bounding boxes"));
        GVDBX_CHECK(hipMemcpyAsync(v->host, v->c_bbox.p, (size_t)nc * sizeof(BBoxAccum), hipMemcpyDeviceToHost, s));
        TRY(sync(v));
        std::vector<BBoxAccum> bbs((size_t)nc);
        memcpy(bbs.data(), v->host, (size_t)nc * sizeof(BBoxAccum));
        // each cloud's level-0 cell box, then shelf packing with 2-cell gaps
        std::vector<int32_t> box((size_t)nc * 6);
        int64_t area = 0;
        int maxx = 1;
        v->pts_valid = 0;
        for (int c = 0; c < nc; c++) {
            CloudFrame f;
            char who[48];
            snprintf(who, sizeof(who), " in cloud %d", c);
            TRY(frame_of(v, bbs[c], radius, who, f));
            v->pts_valid += (int64_t)bbs[c].nvalid;
            for (int a = 0; a < 3; a++) {
                box[6 * c + a] = (int32_t)(f.lo_v[a] / r0);
                box[6 * c + 3 + a] = (int32_t)(f.hi_v[a] / r0) + 1;   // exclusive
            }
            const int sx = box[6 * c + 3] - box[6 * c], sy = box[6 * c + 4] - box[6 * c + 1];
            area += (int64_t)(sx + 2) * (sy + 2);
            maxx = std::max(maxx, sx + 2);
        }
        const int row = std::max<int>(maxx, (int)ceil(sqrt((double)area)));
        v->h_celloff.assign((size_t)nc * 3, 0);
        v->h_boxes.assign((size_t)nc * 6, 0);
        int cx = 2, cy = 2, rowh = 0, ext[3] = {0, 0, 0};
        for (int c = 0; c < nc; c++) {
            const int sx = box[6 * c + 3] - box[6 * c], sy = box[6 * c + 4] - box[6 * c + 1],
                      sz = box[6 * c + 5] - box[6 * c + 2];
            if (cx + sx + 2 > row + 2 && cx > 2) {
                cx = 2;
                cy += rowh + 2;
                rowh = 0;
            }
            const int at[3] = {cx, cy, 2};
            for (int a = 0; a < 3; a++) {
                v->h_celloff[3 * c + a] = at[a] - box[6 * c + a];
                v->h_boxes[6 * c + a] = at[a];
                v->h_boxes[6 * c + 3 + a] = at[a] + (a == 0 ? sx : (a == 1 ? sy : sz));
                ext[a] = std::max(ext[a], v->h_boxes[6 * c + 3 + a] + 2);
            }
            cx += sx + 2;
            rowh = std::max(rowh, sy);
        }
        // the packed space is one volume with its own root
        for (int a = 0; a < 3; a++) {
            pmin[a] = 0.0f;
            pmax[a] = (float)((int64_t)ext[a] * r0);
            lo_v[a] = 0;
            hi_v[a] = ext[a] * r0 - 1;
        }
        depth = determine_depth(v, pmin, pmax, root);
        if (depth <= 0 || depth >= v->levs)
            FAIL(GVDBX_ERR_DEPTH, "RebuildTopology: the packed clouds exceed the configured levels");
        if ((int64_t)ext[0] > 65535 || (int64_t)ext[1] > 65535 || (int64_t)ext[2] > 65535)
            FAIL(GVDBX_ERR_RANGE, "RebuildTopology: the packed clouds exceed 16-bit brick keys");
        TRY(ensure(v, v->c_celloff, (size_t)nc * 12));
        TRY(ensure(v, v->c_boxes, (size_t)nc * 24));
        memcpy(v->host, v->h_celloff.data(), (size_t)nc * 12);
        memcpy((char*)v->host + (size_t)nc * 12, v->h_boxes.data(), (size_t)nc * 24);
        GVDBX_CHECK(hipMemcpyAsync(v->c_celloff.p, v->host, (size_t)nc * 12, hipMemcpyHostToDevice, s));
        GVDBX_CHECK(hipMemcpyAsync(v->c_boxes.p, (char*)v->host + (size_t)nc * 12, (size_t)nc * 24,
                                   hipMemcpyHostToDevice, s));
        TRY(sync(v));
    }

    // validation is done: from here on the previous topology is being replaced
    v->has_topo = false;
    v->has_atlas = false;
    v->topo_gen++;
    v->vis_valid = false;
    int64_t rtop = v->range[depth];
    (void)rtop;
    CellGrid grid[GVDBX_MAXLEV] = {};
    for (int l = 0; l < depth; l++) {
        int64_t rg = v->range[l];
        CellGrid& g = grid[l];
        g.ox = (int)(lo_v[0] / rg); g.oy = (int)(lo_v[1] / rg); g.oz = (int)(lo_v[2] / rg);
        g.nx = (int)(hi_v[0] / rg) - g.ox + 1;
        g.ny = (int)(hi_v[1] / rg) - g.oy + 1;
        g.nz = (int)(hi_v[2] / rg) - g.oz + 1;
    }
    // cell grids: dense over the box, or sparse (the cells the points can
    // reach, found by sorting cell keys) when the box is too large or asked for
    const bool sparse = v->ncloud == 0 && (v->topo_mode == GVDBX_TOPO_RADIX ||
                                           (v->topo_mode == GVDBX_TOPO_AUTO && grid[0].count() > MAX_DENSE_CELLS));
    if (!sparse && grid[0].count() > MAX_DENSE_CELLS)
        FAIL(GVDBX_ERR_RANGE, "RebuildTopology: %lld level-0 cells exceed the dense-grid limit%s",
             (long long)grid[0].count(), v->ncloud ? " (batched clouds)" : " (topology mode DENSE)");
    CellGrid gb_sparse = {};
    if (sparse) TRY(build_sparse_grids(v, n, orig, radius, extra, depth, grid, gb_sparse));
    v->gb_sparse = gb_sparse;
    const int32_t* allowed = nullptr;
    if (v->ncloud > 0) {
        // cell -> cloud over every cloud's cell box; extras only land inside their cloud's box
        const int nc = v->ncloud;
        TRY(ensure(v, v->c_cellcloud, (size_t)grid[0].count() * 4));
        launch_fill_i32(v->c_cellcloud.as<int32_t>(), grid[0].count(), -1, s);
        std::vector<int32_t> tiles;
        for (int c = 0; c < nc; c++) {
            const int32_t* b = &v->h_boxes[6 * c];
            const int64_t cells = (int64_t)(b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
            for (int64_t o = 0; o < cells; o += CELL_CHUNK) {
                tiles.push_back(c);
                tiles.push_back((int32_t)o);
            }
        }
        TRY(ensure(v, v->c_tiles, std::max<size_t>(tiles.size(), 2) * 4));
        TRY(ensure_host(v, tiles.size() * 4 + 4096));
        memcpy(v->host, tiles.data(), tiles.size() * 4);
        GVDBX_CHECK(hipMemcpyAsync(v->c_tiles.p, v->host, tiles.size() * 4, hipMemcpyHostToDevice, s));
        launch_paint_clouds(grid[0], v->c_boxes.as<int32_t>(), v->c_tiles.as<int32_t>(), (int64_t)tiles.size() / 2,
                            v->c_cellcloud.as<int32_t>(), s);
        TRY(check_launch("cloud cells"));
        TRY(sync(v));   // the staging buffer is reused below
        allowed = v->c_cellcloud.as<int32_t>();
    }

    // ---- occupancy ----
    int64_t G0 = grid[0].count();
    TRY(ensure(v, v->reach0, (size_t)G0 * 4));
    for (int l = 0; l < depth; l++) {
        int64_t G = grid[l].count();
        TRY(ensure(v, v->flag_base[l], (size_t)G));
        TRY(ensure(v, v->flag_all[l], (size_t)G));
        TRY(ensure(v, v->idx[l], (size_t)G * 4));
    }
    int64_t ptot = 0;
    int64_t poffs[GVDBX_MAXLEV];
    for (int l = 0; l < depth; l++) {
        poffs[l] = ptot;
        ptot += scan_tiles(grid[l].count()) + 1;
    }
    TRY(ensure(v, v->partials, (size_t)(ptot + 2) * sizeof(U2)));
    TRY(ensure(v, v->totals, (size_t)GVDBX_MAXLEV * sizeof(U2)));
    {
        // reach and level-0 extra flags; the flags of the levels above (level-up only sets them)
        ZeroBatch zb{};
        zb.p[zb.n] = v->reach0.as<uint8_t>(), zb.bytes[zb.n++] = G0 * 4;
        zb.p[zb.n] = v->flag_all[0].as<uint8_t>(), zb.bytes[zb.n++] = G0;
        for (int l = 1; l < depth; l++) {
            zb.p[zb.n] = v->flag_base[l].as<uint8_t>(), zb.bytes[zb.n++] = grid[l].count();
            zb.p[zb.n] = v->flag_all[l].as<uint8_t>(), zb.bytes[zb.n++] = grid[l].count();
        }
        launch_zero_batch(zb, s);
    }
    // bin the points by level-0 cell (kept for InsertPointsSubcell / InsertPoints),
    // then derive occupancy and extra-brick reach from the cell-sorted points
    v->grid[0] = grid[0];
    v->bins_valid = false;
    ReachParams rp{grid[0], r0, v->logdim[0], (extra && mask_ok) ? 1 : 0, radius, cloud_map(v)};
    bool reached = false;
    TRY(ensure_bins(v, n, orig, false, &rp, &reached));
    v->bins_once = true;
    if (!reached)
        launch_reach_sorted(v->sorted.as<float4>(), v->bin_off.as<int32_t>() + v->gbin.count(), n, rp,
                            v->reach0.as<uint32_t>(), s);
    if (extra && !mask_ok)
        launch_mark_boxes(v->pos, n, orig[0], orig[1], orig[2], grid[0], r0, radius, v->flag_all[0].as<uint8_t>(), s);
    launch_expand_reach(grid[0], v->reach0.as<uint32_t>(), v->flag_base[0].as<uint8_t>(), v->flag_all[0].as<uint8_t>(),
                        (extra && !mask_ok) ? 1 : 0, allowed, s);
    for (int l = 1; l < depth; l++)
        launch_level_up(grid[l - 1], v->flag_base[l - 1].as<uint8_t>(), v->flag_all[l - 1].as<uint8_t>(), grid[l],
                        v->logdim[l], v->flag_base[l].as<uint8_t>(), v->flag_all[l].as<uint8_t>(), s);
    U2* partials = v->partials.as<U2>();
    U2* totals = v->totals.as<U2>();
    for (int l = 0; l < depth; l++)
        launch_flag_scan_reduce(v->flag_base[l].as<uint8_t>(), v->flag_all[l].as<uint8_t>(), grid[l].count(),
                                partials + poffs[l], totals + l, s);
    TRY(check_launch("occupancy"));
    GVDBX_CHECK(hipMemcpyAsync(v->host, totals, (size_t)depth * sizeof(U2), hipMemcpyDeviceToHost, s));
    TRY(sync(v));
    U2 tot[GVDBX_MAXLEV];
    memcpy(tot, v->host, (size_t)depth * sizeof(U2));

    // ---- pools (Clear + PoolAlloc, VG.cpp:2487-2503, AL.cpp:163-192) ----
    int64_t cnt[GVDBX_MAXLEV] = {};
    for (int l = 0; l < depth; l++) cnt[l] = (int64_t)tot[l].a + tot[l].b;
    cnt[depth] = 1;
    for (int l = 0; l <= depth; l++) {
        TRY(ensure(v, v->nodes[l], (size_t)cnt[l] * sizeof(GNode)));
        if (l >= 1) {
            int64_t len = (int64_t)v->res[l] * v->res[l] * v->res[l];
            TRY(ensure(v, v->clist[l], (size_t)(cnt[l] * len) * 8));
        }
    }
    FlagApplyBatch fb{};
    for (int l = 0; l < depth; l++) {
        NodeBuildParams np{grid[l], l, v->range[l], l == 0 ? 1 : 0, v->nodes[l].as<GNode>(), v->idx[l].as<int32_t>()};
        fb.lv[fb.n++] = FlagApplyLevel{v->flag_base[l].as<uint8_t>(), v->flag_all[l].as<uint8_t>(), grid[l].count(),
                                       partials + poffs[l], totals + l, np};
    }
    launch_flag_scan_apply(fb, s);
    // root node (ActivateBricksGPU VG.cpp:1290-1301)
    GNode rn;
    memset(&rn, 0, sizeof(rn));
    rn.mLev = (uint8_t)depth;
    rn.mFlags = 1;
    for (int a = 0; a < 3; a++) {
        rn.mPos[a] = root[a];
        rn.mValue[a] = -1;
    }
    rn.mParent = GVDBX_ID_UNDEFL;
    rn.mChildList = 1ull | ((uint64_t)depth << 8);
    // top level (PrepareVDBPartially VG.cpp:3996-4012)
    int tlev = 1;
    for (int l = v->levs - 1; l >= 0; l--)
        if (cnt[l] == 1) tlev = l;
    // root node and every child list ("no child") in one launch
    ClistInit ci{};
    ci.root = rn;
    ci.root_dst = v->nodes[depth].as<GNode>();
    ci.start[0] = 0;
    for (int l = 1; l <= depth; l++) {
        const int64_t len = (int64_t)v->res[l] * v->res[l] * v->res[l];
        ci.list[ci.n] = v->clist[l].as<uint64_t>();
        ci.start[ci.n + 1] = ci.start[ci.n] + cnt[l] * len;
        ci.n++;
    }
    launch_clist_init(ci, s);
    // links (gvdbLinkBricks, levels below top_lev), all levels in one launch
    LinkBatch lb{};
    lb.start[0] = 0;
    for (int l = depth - 1; l >= 0; l--) {
        if (l >= tlev) continue;
        LinkParams lp{};
        lp.lev = l;
        lp.nchild = cnt[l];
        lp.child = v->nodes[l].as<GNode>();
        lp.childw = v->nodes[l].as<GNode>();
        lp.parent_nodes = v->nodes[l + 1].as<GNode>();
        lp.parent_clist = v->clist[l + 1].as<uint64_t>();
        lp.res_p = v->res[l + 1];
        lp.range_p = v->range[l + 1];
        lp.logres_p = v->logdim[l + 1];
        lp.parent_is_root = (l + 1 == depth) ? 1 : 0;
        if (l + 1 < depth) {
            lp.gp = grid[l + 1];
            lp.idx_p = v->idx[l + 1].as<int32_t>();
        }
        lp.nparent = cnt[l + 1];
        lb.lv[lb.n] = lp;
        lb.start[lb.n + 1] = lb.start[lb.n] + cnt[l];
        lb.n++;
    }
    launch_link(lb, s);
    if (v->ncloud > 0) {
        TRY(ensure(v, v->c_leafcloud, (size_t)std::max<int64_t>(cnt[0], 1) * 4));
        launch_leaf_clouds(v->nodes[0].as<GNode>(), cnt[0], grid[0], v->logdim[0], v->c_cellcloud.as<int32_t>(),
                           v->c_leafcloud.as<int32_t>(), s);
    }
    TRY(check_launch("node pools"));

    v->depth = depth;
    v->top_lev = tlev;
    memcpy(v->root_pos, root, sizeof(root));
    memcpy(v->pos_min, pmin, sizeof(pmin));
    memcpy(v->pos_max, pmax, sizeof(pmax));
    for (int l = 0; l < GVDBX_MAXLEV; l++) {
        v->cnt[l] = l <= depth ? cnt[l] : 0;
        v->nbase[l] = l < depth ? (int64_t)tot[l].a : (l == depth ? 1 : 0);
        v->grid[l] = l < depth ? grid[l] : CellGrid{};
    }
    v->has_topo = true;
    v->has_atlas = false;
    v->sc_valid = false;
    v->topo_mode_used = sparse ? GVDBX_TOPO_RADIX : GVDBX_TOPO_DENSE;
    v->leaf_inactive = false;
    v->force_full = false;
    return GVDBX_OK;
}

// ActivateIncreBricksGPU (VG.cpp:1020-1155) on the current tree.  accum:
// AccumulateTopology (VG.cpp:1158-1178), existing leaves keep their flags;
// else the incremental RebuildTopology (VG.cpp:1195-1216), leaves no point
// reaches are deactivated and delinked.  The tree keeps its depth and root.
// With `full` (RebuildTopology), *full is set instead of updating when the
// reference would rebuild (the depth changes) or when a point lies outside
// the root; AccumulateTopology fails with GVDBX_ERR_RANGE there.
int incremental_topology(gvdbx_volume* v, int64_t n, float radius, const float orig[3], bool accum, bool* full) {
    const char* what = accum ? "AccumulateTopology" : "RebuildTopology";
    if (full) *full = false;
    if (n < 0) FAIL(GVDBX_ERR_ARG, "%s: negative point count", what);
    if (!(radius >= 0.0f) || !isfinite(radius)) FAIL(GVDBX_ERR_ARG, "%s: radius must be finite and >= 0", what);
    if (!v->has_pos) FAIL(GVDBX_ERR_STATE, "%s: SetPoints has not been called", what);
    if (!v->has_topo || v->cnt[0] == 0) FAIL(GVDBX_ERR_STATE, "%s: no topology to update", what);
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "%s: no incremental update for batched clouds", what);
    if (v->grid[0].sparse()) {   // the update grows dense grids: a sparse topology is rebuilt in full
        if (full) {
            *full = true;
            return GVDBX_OK;
        }
        FAIL(GVDBX_ERR_STATE, "%s: no incremental update of a sparse (radix-sorted) topology", what);
    }
    if (v->top_lev != v->depth) FAIL(GVDBX_ERR_STATE, "%s: the tree's top level is below its root", what);
    hipStream_t s = v->stream;
    const int depth = v->depth;
    const int64_t rtop = v->range[depth];
    // GetBoundingBox (VG.cpp:5404-5467)
    TRY(ensure(v, v->bbox, sizeof(BBoxAccum) * (1 + (size_t)bbox_parts(std::max<int64_t>(n, 1)))));
    TRY(ensure_host(v, 4096));
    // the bounding-box pass writes every field of the accumulator; only an
    // empty point set needs the initial (empty) box
    if (n > 0) launch_bbox(v->pos, n, orig[0], orig[1], orig[2], v->bbox.as<BBoxAccum>(), v->bbox.as<BBoxAccum>() + 1, s);
    else launch_bbox_init(v->bbox.as<BBoxAccum>(), s);
    TRY(check_launch("bounding box"));
    GVDBX_CHECK(hipMemcpyAsync(v->host, v->bbox.p, sizeof(BBoxAccum), hipMemcpyDeviceToHost, s));
    TRY(sync(v));
    BBoxAccum bb;
    memcpy(&bb, v->host, sizeof(bb));
    if (bb.nvalid == 0) FAIL(GVDBX_ERR_NOPOINTS, "%s: no valid point", what);
    float mn[3], mx[3], pmin[3], pmax[3];
    bool inside = true;
    for (int a = 0; a < 3; a++) {
        memcpy(&mn[a], &bb.mn[a], 4);
        memcpy(&mx[a], &bb.mx[a], 4);
        pmin[a] = std::max(0.0f, mn[a] - 1);
        pmax[a] = std::max(0.0f, mx[a] + 2);
        inside = inside && mn[a] >= (float)v->root_pos[a] && mx[a] < (float)((int64_t)v->root_pos[a] + rtop);
    }
    if (full) {
        int root2[3];
        if (!inside || determine_depth(v, pmin, pmax, root2) != depth) {
            *full = true;
            return GVDBX_OK;
        }
    } else if (!inside) {
        FAIL(GVDBX_ERR_RANGE, "%s: points lie outside the root node at (%d,%d,%d), %lld voxels wide", what,
             v->root_pos[0], v->root_pos[1], v->root_pos[2], (long long)rtop);
    }
    // grids that hold the old ones and every cell a box can touch (the box
    // corners are monotone in the point coordinate)
    CellGrid ug[GVDBX_MAXLEV] = {};
    IncreMark im{};
    im.depth = depth;
    for (int l = 0; l < depth; l++) {
        const int64_t rg = v->range[l];
        const CellGrid& og = v->grid[l];
        int lo[3], hi[3];
        im.range[l] = rg;
        im.root_span[l] = (int)(rtop / rg);
        for (int a = 0; a < 3; a++) {
            im.root_cell[l][a] = (int)(v->root_pos[a] / rg);
            const int bmin = (int)((float)(int)mn[a] - radius), bmax = (int)((float)(int)mx[a] + radius);
            lo[a] = std::max((int)(bmin / rg), im.root_cell[l][a]);
            hi[a] = std::min((int)(bmax / rg), im.root_cell[l][a] + im.root_span[l] - 1);
        }
        const int olo[3] = {og.ox, og.oy, og.oz}, ohi[3] = {og.ox + og.nx - 1, og.oy + og.ny - 1, og.oz + og.nz - 1};
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], olo[a]);
            hi[a] = std::max(hi[a], ohi[a]);
        }
        ug[l] = CellGrid{lo[0], lo[1], lo[2], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    }
    if (ug[0].count() > MAX_DENSE_CELLS) {
        if (full) {
            *full = true;
            return GVDBX_OK;
        }
        FAIL(GVDBX_ERR_RANGE, "%s: %lld level-0 cells exceed the dense-grid limit", what, (long long)ug[0].count());
    }
    const int64_t nleaf = v->cnt[0];
    int64_t ptot = 0, poffs[GVDBX_MAXLEV];
    for (int l = 0; l < depth; l++) {
        TRY(ensure(v, v->idx2[l], (size_t)ug[l].count() * 4));
        TRY(ensure(v, v->flag_base[l], (size_t)ug[l].count()));
        poffs[l] = ptot;
        ptot += scan_tiles(ug[l].count()) + 1;
    }
    TRY(ensure(v, v->inc_mark, (size_t)nleaf));
    TRY(ensure(v, v->partials, (size_t)(ptot + 2) * sizeof(U2)));
    TRY(ensure(v, v->totals, (size_t)GVDBX_MAXLEV * sizeof(U2)));
    for (int l = 0; l < depth; l++) {
        launch_remap_idx(v->grid[l], v->idx[l].as<int32_t>(), ug[l], v->idx2[l].as<int32_t>(), s);
        launch_fill_u8(v->flag_base[l].as<uint8_t>(), ug[l].count(), 0, s);
        im.g[l] = ug[l];
        im.idx[l] = v->idx2[l].as<int32_t>();
        im.newf[l] = v->flag_base[l].as<uint8_t>();
    }
    launch_fill_u8(v->inc_mark.as<uint8_t>(), nleaf, 0, s);
    im.marker0 = v->inc_mark.as<uint8_t>();
    im.nleaf = nleaf;
    launch_incre_mark(v->pos, n, orig[0], orig[1], orig[2], radius, im, s);
    U2* partials = v->partials.as<U2>();
    U2* totals = v->totals.as<U2>();
    for (int l = 0; l < depth; l++)
        launch_flag_scan_reduce(v->flag_base[l].as<uint8_t>(), v->flag_base[l].as<uint8_t>(), ug[l].count(),
                                partials + poffs[l], totals + l, s);
    TRY(check_launch("incremental marks"));
    GVDBX_CHECK(hipMemcpyAsync(v->host, totals, (size_t)depth * sizeof(U2), hipMemcpyDeviceToHost, s));
    TRY(sync(v));
    U2 tot[GVDBX_MAXLEV];
    memcpy(tot, v->host, (size_t)depth * sizeof(U2));
    memcpy(v->pos_min, pmin, sizeof(pmin));
    memcpy(v->pos_max, pmax, sizeof(pmax));
    v->pts_valid = (int64_t)bb.nvalid;
    v->bins_valid = false;   // the bins hold earlier points
    v->sc_valid = false;
    int64_t added = 0;
    for (int l = 0; l < depth; l++) added += tot[l].a;
    if (added == 0) return GVDBX_OK;   // no missing node: nothing changes (VG.cpp:1066)

    // pools grow by the new nodes, keeping the old records and lists
    int64_t cnt[GVDBX_MAXLEV];
    for (int l = 0; l < depth; l++) {
        cnt[l] = v->cnt[l] + (int64_t)tot[l].a;
        TRY(ensure_keep(v, v->nodes[l], (size_t)cnt[l] * sizeof(GNode), (size_t)v->cnt[l] * sizeof(GNode)));
        if (l >= 1) {
            const int64_t len = (int64_t)v->res[l] * v->res[l] * v->res[l];
            TRY(ensure_keep(v, v->clist[l], (size_t)(cnt[l] * len) * 8, (size_t)(v->cnt[l] * len) * 8));
            launch_fill_u64(v->clist[l].as<uint64_t>() + v->cnt[l] * len, (cnt[l] - v->cnt[l]) * len,
                            GVDBX_ID_UNDEF64, s);
        }
    }
    // leaf flags from the markers (VG.cpp:1078-1085)
    launch_incre_leaf_flags(v->nodes[0].as<GNode>(), nleaf, v->inc_mark.as<uint8_t>(), accum ? 1 : 0, s);
    // new nodes after the old ones, in key order, entered in the maps (VG.cpp:1087-1111)
    FlagApplyBatch fb{};
    for (int l = 0; l < depth; l++) {
        NodeBuildParams np{ug[l], l, v->range[l], l == 0 ? 1 : 0, v->nodes[l].as<GNode>(), v->idx2[l].as<int32_t>(),
                           v->cnt[l], 1};
        fb.lv[fb.n++] = FlagApplyLevel{v->flag_base[l].as<uint8_t>(), v->flag_base[l].as<uint8_t>(), ug[l].count(),
                                       partials + poffs[l], totals + l, np};
    }
    launch_flag_scan_apply(fb, s);
    // links of the flagged nodes (gvdbLinkBricks; the maps give what a descent finds)
    LinkBatch lb{};
    lb.start[0] = 0;
    for (int l = depth - 1; l >= 0; l--) {
        LinkParams lp{};
        lp.lev = l;
        lp.nchild = cnt[l];
        lp.child = v->nodes[l].as<GNode>();
        lp.childw = v->nodes[l].as<GNode>();
        lp.parent_nodes = v->nodes[l + 1].as<GNode>();
        lp.parent_clist = v->clist[l + 1].as<uint64_t>();
        lp.res_p = v->res[l + 1];
        lp.range_p = v->range[l + 1];
        lp.logres_p = v->logdim[l + 1];
        lp.parent_is_root = (l + 1 == depth) ? 1 : 0;
        if (l + 1 < depth) {
            lp.gp = ug[l + 1];
            lp.idx_p = v->idx2[l + 1].as<int32_t>();
        }
        lp.nparent = l + 1 < depth ? cnt[l + 1] : 1;
        lp.need_flag = 1;
        lb.lv[lb.n] = lp;
        lb.start[lb.n + 1] = lb.start[lb.n] + cnt[l];
        lb.n++;
    }
    launch_link(lb, s);
    // delinks, leaves first, then one inner level after the other (ND.cuh:401-467)
    for (int l = 0; l < depth; l++) {
        DelinkParams dp{};
        dp.lev = l;
        dp.nodes = v->nodes[l].as<GNode>();
        dp.n = cnt[l];
        if (l >= 1) {
            dp.own_clist = v->clist[l].as<uint64_t>();
            dp.own_len = (int64_t)v->res[l] * v->res[l] * v->res[l];
        }
        dp.parent_is_root = (l + 1 == depth) ? 1 : 0;
        if (l + 1 < depth) {
            dp.gp = ug[l + 1];
            dp.idx_p = v->idx2[l + 1].as<int32_t>();
        }
        dp.parent_nodes = v->nodes[l + 1].as<GNode>();
        dp.parent_clist = v->clist[l + 1].as<uint64_t>();
        dp.res_p = v->res[l + 1];
        dp.range_p = v->range[l + 1];
        dp.nparent = l + 1 < depth ? cnt[l + 1] : 1;
        launch_incre_delink(dp, s);
    }
    for (int l = 0; l < depth; l++) launch_incre_final_idx(ug[l], v->idx2[l].as<int32_t>(), v->nodes[l].as<GNode>(), s);
    TRY(check_launch("incremental topology"));
    for (int l = 0; l < depth; l++) {
        std::swap(v->idx[l], v->idx2[l]);
        v->grid[l] = ug[l];
        v->cnt[l] = cnt[l];
    }
    v->topo_gen++;
    v->has_atlas = false;
    v->vis_valid = false;
    v->leaf_inactive = true;
    return GVDBX_OK;
}

int compute_bounds(gvdbx_volume* v) {
    // ComputeBounds VG.cpp:1792-1816
    hipStream_t s = v->stream;
    TRY(ensure(v, v->misc, 256));
    TRY(ensure_pin(v));
    // an unread earlier readback into the same slot is superseded: the host
    // never writes the slot, and stream order lands this copy last
    launch_bounds_init(v->misc.as<int32_t>(), s);
    launch_leaf_bounds(v->nodes[0].as<GNode>(), v->cnt[0], v->res[0], v->misc.as<int32_t>(), s);
    TRY(check_launch("leaf bounds"));
    GVDBX_CHECK(hipMemcpyAsync((char*)v->pin + PIN_BOUNDS, v->misc.p, 6 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    v->bounds_pending = true;
    if (v->ncloud > 0) {
        // ComputeBounds of every cloud, in its own coordinates
        const int nc = v->ncloud;
        TRY(ensure(v, v->c_bounds, (size_t)nc * 24));
        TRY(ensure_host(v, (size_t)nc * 24 + 4096));
        std::vector<int32_t> init6((size_t)nc * 6);
        for (int c = 0; c < nc; c++)
            for (int a = 0; a < 3; a++) {
                init6[6 * c + a] = 0x7FFFFFFF;
                init6[6 * c + 3 + a] = INT32_MIN;
            }
        memcpy(v->host, init6.data(), (size_t)nc * 24);
        GVDBX_CHECK(hipMemcpyAsync(v->c_bounds.p, v->host, (size_t)nc * 24, hipMemcpyHostToDevice, s));
        launch_leaf_bounds_clouds(v->nodes[0].as<GNode>(), v->cnt[0], v->res[0], v->c_leafcloud.as<int32_t>(),
                                  v->c_bounds.as<int32_t>(), s);
        TRY(check_launch("cloud bounds"));
        GVDBX_CHECK(hipMemcpyAsync(v->host, v->c_bounds.p, (size_t)nc * 24, hipMemcpyDeviceToHost, s));
        TRY(sync(v));
        memcpy(init6.data(), v->host, (size_t)nc * 24);
        v->h_cbounds.assign((size_t)nc * 6, 0.0f);
        for (int c = 0; c < nc; c++)
            for (int a = 0; a < 3; a++) {
                const int32_t sh = v->h_celloff[3 * c + a] * v->res[0];
                v->h_cbounds[6 * c + a] = (float)(init6[6 * c + a] - sh);
                v->h_cbounds[6 * c + 3 + a] = (float)(init6[6 * c + 3 + a] - sh);
            }
    }
    return GVDBX_OK;
}

int update_atlas(gvdbx_volume* v) {
    if (!v->has_topo) FAIL(GVDBX_ERR_STATE, "UpdateAtlas: no topology");
    if (!any_chan(v)) FAIL(GVDBX_ERR_STATE, "UpdateAtlas: no channel (AddChannel)");
    if (v->leaf_inactive) {
        // after an incremental update only the active leaves take atlas
        // slots, in leaf order (VG.cpp:2654-2661); each leaf's voxels stay at
        // its own index in the device buffers, so those hold every leaf
        hipStream_t s = v->stream;
        const int64_t nl = v->cnt[0];
        TRY(ensure(v, v->atl_flag, (size_t)nl * 4));
        TRY(ensure(v, v->atl_rank, (size_t)(nl + 1) * 4));
        TRY(ensure(v, v->partials, (size_t)(scan_tiles(nl + 1) + 2) * sizeof(U2)));
        TRY(ensure(v, v->totals, (size_t)GVDBX_MAXLEV * sizeof(U2)));
        TRY(ensure_host(v, 4096));
        launch_leaf_flags32(v->nodes[0].as<GNode>(), nl, v->atl_flag.as<int32_t>(), s);
        launch_count_scan_reduce(v->atl_flag.as<int32_t>(), nl, v->partials.as<uint32_t>(), v->totals.as<uint32_t>(), s);
        launch_count_scan_apply(v->atl_flag.as<int32_t>(), nl, v->partials.as<uint32_t>(), v->atl_rank.as<int32_t>(), s);
        TRY(check_launch("atlas ranks"));
        GVDBX_CHECK(hipMemcpyAsync(v->host, v->atl_rank.as<int32_t>() + nl, 4, hipMemcpyDeviceToHost, s));
        TRY(sync(v));
        int32_t used = 0;
        memcpy(&used, v->host, 4);
        int64_t amax = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
        if (used > amax) {   // VG.cpp:2645-2651 + AtlasResize AL.cpp:622-650
            int sz = (int)ceil(used / (float)(v->subdim[0] * v->subdim[1]));
            if (sz > v->subdim[2]) v->subdim[2] = sz;
        }
        int64_t slots = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
        TRY(grow_atlas(v, std::max(slots, nl)));
        TRY(ensure(v, v->amap, (size_t)slots * sizeof(GAtlasNode)));
        launch_atlas_assign_rank(v->nodes[0].as<GNode>(), nl, v->atl_rank.as<int32_t>(), v->subdim[0], v->subdim[1],
                                 brick_res(v), std::max(v->apron, 0), v->amap.as<GAtlasNode>(), slots, s);
        TRY(check_launch("atlas assign"));
        v->amap_cnt = slots;
        v->has_atlas = true;
        return GVDBX_OK;
    }
    int64_t used = v->cnt[0];
    int64_t amax = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
    if (used > amax) {   // VG.cpp:2645-2651 + AtlasResize AL.cpp:622-650
        int sz = (int)ceil(used / (float)(v->subdim[0] * v->subdim[1]));
        if (sz > v->subdim[2]) v->subdim[2] = sz;
    }
    int64_t slots = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
    TRY(grow_atlas(v, slots));
    TRY(ensure(v, v->amap, (size_t)slots * sizeof(GAtlasNode)));
    launch_atlas_assign(v->nodes[0].as<GNode>(), v->cnt[0], v->subdim[0], v->subdim[1], brick_res(v),
                        std::max(v->apron, 0), v->amap.as<GAtlasNode>(), slots, v->stream);
    TRY(check_launch("atlas assign"));
    v->amap_cnt = slots;
    v->has_atlas = true;
    return GVDBX_OK;
}

int need_atlas(gvdbx_volume* v, const char* what) {
    if (!v->has_topo) FAIL(GVDBX_ERR_STATE, "%s: no topology", what);
    if (!v->has_atlas) FAIL(GVDBX_ERR_STATE, "%s: UpdateAtlas has not been called", what);
    if (v->cnt[0] == 0) FAIL(GVDBX_ERR_STATE, "%s: no bricks", what);
    return GVDBX_OK;
}

int gather_color(gvdbx_volume* v, int chan_clr, int accum) {
    ColorGatherParams cp{};
    cp.leaves = v->nodes[0].as<GNode>();
    cp.nleaf = v->cnt[0];
    cp.gb = v->gbin;
    cp.off = v->bin_off.as<int32_t>();
    cp.sorted = v->sorted.as<float4>();
    cp.clr = v->clr;
    cp.clr_stride = v->clr_stride;
    cp.chan = v->chan[chan_clr].as<uint32_t>();
    cp.brick_res = brick_res(v);
    cp.apron = std::max(v->apron, 0);
    cp.res0 = v->res[0];
    cp.logr0 = v->logdim[0];
    cp.sc = v->sc_size;
    cp.sc_shift = __builtin_ctz((unsigned)v->sc_size);
    cp.R = v->sc_radius;
    cp.accum = accum ? 1 : 0;
    chan_written(v, chan_clr);
    launch_gather_color(cp, v->stream);
    return check_launch("gather colour");
}

// noalloc: the packed ranges go to the library's static scratch, and no
// other scratch is used (no reach classes, no split tiles), so the call
// allocates nothing; the result is the same bit for bit.
int gather(gvdbx_volume* v, int mode, int sc, float radius, int64_t sc_len, int chan, int chan_clr, int accum,
           bool noalloc = false, bool fp16 = false) {
    const char* what = mode ? "GatherLevelSet" : "GatherDensity";
    if (noalloc) what = "GatherLevelSetNoAlloc";
    if (fp16) what = "GatherLevelSet_FP16";
    if (v->sc_valid && v->sc_fp16 != fp16)
        FAIL(GVDBX_ERR_STATE, "%s: the subcell lists were made by InsertPointsSubcell%s", what, v->sc_fp16 ? "_FP16" : "");
    if (fp16 && v->clr)
        FAIL(GVDBX_ERR_ARG, "%s: colours are not defined for FP16 lists (the reference sums an uninitialised colour)",
             what);
    TRY(need_atlas(v, what));
    TRY(check_chan_type(v, chan, GVDBX_FLOAT, what));
    // point colours make the gather average them into chan_clr
    // (cuda_gvdb_particles.cuh:718-731,757-771: active when the subcell
    // colour list exists, i.e. SetPoints was given colours)
    if (v->clr) TRY(check_chan_type(v, chan_clr, GVDBX_UCHAR4, what));
    if (!v->sc_valid || !v->bins_valid) FAIL(GVDBX_ERR_STATE, "%s: InsertPointsSubcell has not been called", what);
    if (sc != v->sc_size) FAIL(GVDBX_ERR_ARG, "%s: subcell size %d differs from InsertPointsSubcell (%d)", what, sc, v->sc_size);
    if (!(radius >= 0.0f) || !isfinite(radius)) FAIL(GVDBX_ERR_ARG, "%s: bad radius", what);
    if (mode == 0 && sc_len == 0) return GVDBX_OK;   // VG.cpp:5321
    if (v->clr) TRY(gather_color(v, chan_clr, accum));
    GatherParams gp{};
    gp.leaves = v->nodes[0].as<GNode>();
    gp.nleaf = v->cnt[0];
    gp.gb = v->gbin;
    gp.off = v->bin_off.as<int32_t>();
    gp.sorted = v->sorted.as<float4>();
    gp.chan = v->chan[chan].as<float>();
    gp.brick_res = brick_res(v);
    gp.apron = std::max(v->apron, 0);
    gp.mode = mode;
    gp.accum = accum ? 1 : 0;
    gp.sc = sc;
    gp.sc_shift = __builtin_ctz((unsigned)sc);
    gp.box_radius = v->sc_radius;
    gp.radius = radius;
    // Points farther than the cutoff cannot change a voxel: the level set
    // starts at 3.0 and a member at distance c contributes c - r, the density
    // starts at 0 and a member contributes r - c.  The margin covers FP32
    // rounding of the distance (relative error < 1e-6).
    gp.cutoff = accum ? 1.0e7f : (mode == 1 ? 3.0f + radius + 1.0e-3f : radius + 1.0e-3f);
    gp.variant = (noalloc || fp16) ? 1 : v->gather_variant;
    gp.xcd_remap = v->xcd_remap;
    gp.quarters = v->gather_quarters;
    // 8^3 bricks with apron 1: the walk's store also fills the neighbours' aprons
    // (not for the allocation-free call: its leaf table would be one more buffer)
    gp.push = (v->apron_push && gp.variant == 1 && gp.brick_res == 10 && gp.apron == 1 && v->res[0] == 8 &&
               !v->chan_raw[chan] && !noalloc) ? 1 : 0;
    if (gp.push) {
        TRY(ensure(v, v->bin_leaf, (size_t)std::max<int64_t>(gp.gb.count(), 1) * 4));
        launch_bin_leaf(gp.gb, v->grid[0], v->idx[0].as<int32_t>(), v->cnt[0], v->bin_leaf.as<int32_t>(), v->stream);
        gp.bin_leaf = v->bin_leaf.as<int32_t>();
    }
    v->apron_pushed[chan] = gp.push != 0;
    v->apron_push_gen[chan] = v->topo_gen;
    if (fp16 && (int)std::ceil(gp.box_radius) + sc > 28)
        FAIL(GVDBX_ERR_RANGE, "%s: subcell radius + subcell size beyond 28 voxels", what);
    if (gp.variant == 1 && (int)std::ceil(gp.box_radius) + sc <= 29 && v->bins_n > 0) {
        // each point's clipped ranges once, not once per tile that reads it
        if (noalloc) {
            if ((uint64_t)v->bins_n * 4 > GVDBX_NOALLOC_SCRATCH_BYTES)
                FAIL(GVDBX_ERR_RANGE, "%s: %lld points need %lld bytes of scratch, more than its %llu", what,
                     (long long)v->bins_n, (long long)v->bins_n * 4, (unsigned long long)GVDBX_NOALLOC_SCRATCH_BYTES);
            uint32_t* box = nullptr;
            TRY(scratch_on(v->device, &box));
            launch_gather_box(gp.sorted, gp.off + gp.gb.count(), v->bins_n, gp.box_radius, sc, gp.sc_shift, gp.cutoff,
                              box, v->stream);
            gp.box = box;
        } else if (v->gather_classes || fp16) {
            // and each bin's entries grouped by the neighbour cells they can
            // reach, so a tile reads only the groups that can reach it
            const int64_t nbin = gp.gb.count();
            TRY(ensure(v, v->gcls_pts, (size_t)v->bins_n * 16));
            TRY(ensure(v, v->gcoff, (size_t)(nbin * 8 + 1) * 4));
            TRY(ensure(v, v->gcboth, (size_t)std::max<int64_t>(nbin, 1)));
            // bins of more than 256 entries: at most entries / 256 of them, cut
            // into at most entries / 256 + (their number) units of 256
            const int64_t big_cap = v->bins_n / 256 + 1, unit_cap = v->bins_n / 128 + 2;
            TRY(ensure(v, v->gcctr, 8));
            TRY(ensure(v, v->gcunit, (size_t)unit_cap * 8));
            TRY(ensure(v, v->gcbig, (size_t)big_cap * 16));
            TRY(ensure(v, v->gcucnt, (size_t)unit_cap * 36));
            launch_gather_classes(gp.sorted, gp.off, gp.gb, gp.brick_res - 2 * gp.apron, gp.box_radius, sc, gp.sc_shift,
                                  gp.cutoff, v->gcls_pts.as<float4>(), v->gcoff.as<int32_t>(), v->gcboth.as<uint8_t>(),
                                  v->gcctr.as<int32_t>(), v->gcunit.as<int2>(), unit_cap, v->gcbig.as<int4>(), big_cap,
                                  v->gcucnt.as<int32_t>(), v->stream, fp16 ? &v->sc_q : nullptr);
            gp.sorted = v->gcls_pts.as<float4>();
            gp.coff = v->gcoff.as<int32_t>();
            gp.cboth = v->gcboth.as<uint8_t>();
        } else {
            TRY(ensure(v, v->gbox, (size_t)v->bins_n * 4));
            launch_gather_box(gp.sorted, gp.off + gp.gb.count(), v->bins_n, gp.box_radius, sc, gp.sc_shift, gp.cutoff,
                              v->gbox.as<uint32_t>(), v->stream);
            gp.box = v->gbox.as<uint32_t>();
        }
    }
    if (gp.variant == 1 && v->gather_split > 0 && !noalloc) {
        // Bounds: a bin is a neighbour of at most 27 leaves, so the chunks
        // summed over all tiles are at most 27 * tiles * (chunks of all bins),
        // and those are at most ceil(n/64) + non-empty bins (<= min(n, bins)).
        const int res0 = gp.brick_res - 2 * gp.apron;
        const int td = std::min(res0, 8);
        const int64_t tpb = (int64_t)(res0 / td) * (res0 / td) * (res0 / td);
        const int64_t ntiles = gp.nleaf * tpb;
        const int64_t nb = std::max<int64_t>(v->bins_n, 0);
        const double chunks = 27.0 * (double)tpb * ((double)((nb + 63) / 64) + (double)std::min<int64_t>(nb, gp.gb.count()));
        const double split = (double)v->gather_split;
        const double item_cap = std::ceil(chunks / split) + 64.0;   // items beyond the first of each tile
        const int64_t slot_cap = std::min<int64_t>(ntiles, (int64_t)std::ceil(chunks / split) + 1);
        if (item_cap < (double)INT_MAX && ntiles < INT_MAX) {
            TRY(ensure(v, v->gs_items, (size_t)item_cap * 16));
            TRY(ensure(v, v->gs_counts, 8));
            TRY(ensure(v, v->gs_slots, (size_t)std::max<int64_t>(slot_cap, 1) * 4));
            TRY(ensure(v, v->gs_acc, (size_t)std::max<int64_t>(slot_cap, 1) * 2048));
            launch_gather_split(gp, v->gather_split, v->gs_items.as<int4>(), (int64_t)item_cap,
                                v->gs_counts.as<int32_t>(), v->gs_slots.as<int32_t>(), slot_cap,
                                v->gs_acc.as<uint32_t>(), v->stream);
            return check_launch(what);
        }
    }
    launch_gather(gp, v->stream);
    return check_launch(what);
}

int bits_for(int64_t n) {   // bits to hold values 0..n-1 (at least 1)
    int b = 1;
    while (b < 62 && ((int64_t)1 << b) < n) b++;
    return b;
}

void fill_cam(const gvdbx_camera* cam, RenderCam& rc) {
    memset(&rc, 0, sizeof(rc));
    rc.width = cam->width;
    rc.height = cam->height;
    memcpy(rc.org, cam->org, sizeof(rc.org));
    memcpy(rc.org_u, cam->org_u, sizeof(rc.org_u));
    memcpy(rc.org_v, cam->org_v, sizeof(rc.org_v));
    memcpy(rc.dir, cam->dir, sizeof(rc.dir));
    memcpy(rc.dir_u, cam->dir_u, sizeof(rc.dir_u));
    memcpy(rc.dir_v, cam->dir_v, sizeof(rc.dir_v));
}

int make_render(gvdbx_volume* v, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp, const char* what,
                RenderTree& tv, RenderCam& rc, RenderScene& sc) {
    TRY(need_atlas(v, what));
    TRY(check_chan_type(v, chan, GVDBX_FLOAT, what));
    if (!cam || !rp) FAIL(GVDBX_ERR_ARG, "%s: null camera or parameters", what);
    if (v->apron != 1) FAIL(GVDBX_ERR_ARG, "%s: rendering needs apron 1 (AddChannel)", what);
    if (cam->width <= 0 || cam->height <= 0 || (int64_t)cam->width * cam->height >= ((int64_t)1 << 31))
        FAIL(GVDBX_ERR_ARG, "%s: bad image size %dx%d", what, cam->width, cam->height);
    if (!(rp->step > 0.0f) || !isfinite(rp->step)) FAIL(GVDBX_ERR_ARG, "%s: step must be > 0", what);
    if (rp->max_iter < 0) FAIL(GVDBX_ERR_ARG, "%s: negative max_iter", what);
    memset(&tv, 0, sizeof(tv));
    for (int l = 0; l < v->levs && l < GVDBX_MAXLEV; l++) {
        tv.nodes[l] = v->nodes[l].as<GNode>();
        tv.clist[l] = v->clist[l].as<uint64_t>();
        tv.ncnt[l] = v->cnt[l];
        tv.logdim[l] = v->logdim[l];
        tv.res[l] = v->res[l];
        tv.vdel[l] = l > 0 ? (float)v->range[l - 1] : 1.0f;
    }
    tv.levs = v->levs;
    tv.top_lev = v->top_lev;
    TRY(resolve_bounds(v));
    memcpy(tv.bmin, v->obj_min, sizeof(tv.bmin));
    memcpy(tv.bmax, v->obj_max, sizeof(tv.bmax));
    tv.chan = v->chan[chan].as<float>();
    tv.brick_res = brick_res(v);
    memset(&rc, 0, sizeof(rc));
    rc.width = cam->width;
    rc.height = cam->height;
    memcpy(rc.org, cam->org, sizeof(rc.org));
    memcpy(rc.org_u, cam->org_u, sizeof(rc.org_u));
    memcpy(rc.org_v, cam->org_v, sizeof(rc.org_v));
    memcpy(rc.dir, cam->dir, sizeof(rc.dir));
    memcpy(rc.dir_u, cam->dir_u, sizeof(rc.dir_u));
    memcpy(rc.dir_v, cam->dir_v, sizeof(rc.dir_v));
    sc.step = rp->step;
    sc.extinct = rp->extinct;
    sc.albedo = rp->albedo;
    sc.minval = rp->minval;
    sc.alphacut = rp->alphacut;
    sc.thresh = rp->thresh;
    sc.vmin = rp->vmin;
    sc.vmax = rp->vmax;
    sc.epsilon = rp->epsilon;
    sc.max_iter = rp->max_iter;
    memcpy(sc.backclr, rp->backclr, sizeof(sc.backclr));
    memcpy(sc.light, rp->light, sizeof(sc.light));
    sc.shadow_amt = rp->shadow_amt;
    sc.shadow_bias = rp->shadow_bias;
    sc.lut = (const float4*)rp->lut_dev;
    sc.finestep = rp->finestep;
    memcpy(sc.slice_pnt, rp->slice_pnt, sizeof(sc.slice_pnt));
    memcpy(sc.slice_norm, rp->slice_norm, sizeof(sc.slice_norm));
    sc.clr = nullptr;   // Render sets it (Raytrace and the EA raycast take no colour)
    return GVDBX_OK;
}

}  // namespace

// =====================================================================
// C-ABI
// =====================================================================
extern "C" {

const char* gvdbx_last_error(void) { return g_err; }
const char* gvdbx_version(void) { return "gvdbx 0.1 (gfx950)"; }

int gvdbx_create(gvdbx_volume** out, int device) {
    if (!out) FAIL(GVDBX_ERR_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) FAIL(GVDBX_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) FAIL(GVDBX_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    gvdbx_volume* v = new gvdbx_volume();
    v->device = device;
    int st = set_dev(v);
    uint32_t* sp = nullptr;
    if (st == GVDBX_OK) st = scratch_on(device, &sp);   // set-up: the no-alloc scratch's address on this device
    if (st != GVDBX_OK) {
        delete v;
        return st;
    }
    *out = v;
    return GVDBX_OK;
}

int gvdbx_destroy(gvdbx_volume* v) {
    if (!v) FAIL(GVDBX_ERR_ARG, "null volume handle");
    (void)hipSetDevice(v->device);
    (void)hipStreamSynchronize(v->stream);
    retire_graphs(v);
    for (int c = 0; c < GVDBX_MAXCHAN; c++) free_buf(v->chan[c]);
    for (int l = 0; l <= GVDBX_MAXLEV; l++) {
        free_buf(v->nodes[l]);
        free_buf(v->clist[l]);
    }
    for (int l = 0; l < GVDBX_MAXLEV; l++) {
        free_buf(v->idx[l]);
        free_buf(v->idx2[l]);
        free_buf(v->flag_base[l]);
        free_buf(v->flag_all[l]);
        free_buf(v->sg_keys[l]);
        free_buf(v->sg_ht[l]);
    }
    DevBuf* bufs[] = {&v->sg_bkeys, &v->sg_bht, &v->sg_k1, &v->sg_k2, &v->sg_v1, &v->sg_v2, &v->sg_flag, &v->sg_off, &v->sg_part,
                      &v->inc_mark, &v->atl_flag, &v->atl_rank, &v->reach0, &v->nbmask, &v->partials, &v->totals, &v->bbox, &v->misc, &v->amap,
                      &v->bin_off, &v->sorted, &v->sorted2, &v->bkey, &v->bval, &v->bkey2, &v->bval2, &v->rs_hist,
                      &v->rs_off, &v->rs_part, &v->pnode, &v->poff,
                      &v->vis_key, &v->vis_rec, &v->vis_idx, &v->vis_tkey, &v->vis_tidx,
                      &v->vis_hist, &v->vis_off, &v->vis_part, &v->vis_seg, &v->vis_cnt,
                      &v->splat_src, &v->gbox, &v->bin_leaf, &v->gcls_pts, &v->gcoff, &v->gcboth, &v->gcctr, &v->gcunit, &v->gcbig, &v->gcucnt, &v->gs_items, &v->gs_counts, &v->gs_slots, &v->gs_acc,
                      &v->sp_runs, &v->sp_runoff, &v->sp_runcell, &v->sp_part, &v->sp_scan,
                      &v->c_ptoff, &v->c_celloff, &v->c_boxes, &v->c_tiles, &v->c_cellcloud, &v->c_leafcloud,
                      &v->c_bbox, &v->c_bounds, &v->c_cams, &v->vis_lkey, &v->vis_ctag,
                      &v->vis_cbase, &v->vis_csize, &v->vis_coff};
    for (DevBuf* b : bufs) free_buf(*b);
    if (v->host) (void)hipHostFree(v->host);
    if (v->pin) (void)hipHostFree(v->pin);
    delete v;
    return GVDBX_OK;
}

int gvdbx_set_stream(gvdbx_volume* v, void* stream) {
    TRY(check_vol(v));
    v->stream = (hipStream_t)stream;
    return GVDBX_OK;
}

int gvdbx_synchronize(gvdbx_volume* v) {
    TRY(check_vol(v));
    return sync(v);
}

int gvdbx_configure(gvdbx_volume* v, int levs, const int* log2dims) {
    TRY(check_vol(v));
    if (levs < 2 || levs > GVDBX_MAXLEV || !log2dims) FAIL(GVDBX_ERR_ARG, "Configure: levels must be in [2,%d]", GVDBX_MAXLEV);
    int ld[GVDBX_MAXLEV];
    int64_t rg = 1;
    for (int l = 0; l < levs; l++) {
        ld[l] = log2dims[l] == 0 ? 1 : log2dims[l];   // VG.cpp:2394
        if (ld[l] < 1 || ld[l] > 10) FAIL(GVDBX_ERR_ARG, "Configure: log2 dim %d at level %d out of range", ld[l], l);
        rg *= (int64_t)1 << ld[l];
        if (rg > (int64_t)1 << 30) FAIL(GVDBX_ERR_ARG, "Configure: tree range exceeds 2^30 voxels");
    }
    if ((1 << ld[0]) > 32) FAIL(GVDBX_ERR_ARG, "Configure: bricks larger than 32^3 are not supported");
    v->levs = levs;
    for (int l = 0; l < GVDBX_MAXLEV; l++) {
        v->logdim[l] = l < levs ? ld[l] : 0;
        v->res[l] = l < levs ? 1 << ld[l] : 0;
    }
    for (int l = 0; l < levs; l++) v->range[l] = l == 0 ? v->res[0] : v->range[l - 1] * v->res[l];
    v->configured = true;
    v->has_topo = false;
    v->has_atlas = false;
    v->bins_valid = false;
    v->sc_valid = false;
    return GVDBX_OK;
}

int gvdbx_set_channel_default(gvdbx_volume* v, int sx, int sy, int sz) {
    TRY(check_vol(v));
    if (sx < 1 || sy < 1 || sz < 1) FAIL(GVDBX_ERR_ARG, "SetChannelDefault: counts must be >= 1");
    v->def_axis[0] = sx;
    v->def_axis[1] = sy;
    v->def_axis[2] = sz;
    return GVDBX_OK;
}

int gvdbx_set_color_channel(gvdbx_volume* v, int chan) {
    TRY(check_vol(v));
    // the reference stores a uchar: -1 and 255 (CHAN_UNDEF) both mean none
    if (chan == -1 || chan == 255) {
        v->clr_chan = -1;
        return GVDBX_OK;
    }
    if (chan < 0 || chan >= GVDBX_MAXCHAN) FAIL(GVDBX_ERR_ARG, "SetColorChannel: channel %d out of range", chan);
    v->clr_chan = chan;
    return GVDBX_OK;
}

int gvdbx_get_color_channel(gvdbx_volume* v, int* chan) {
    TRY(check_vol(v));
    if (!chan) FAIL(GVDBX_ERR_ARG, "GetColorChannel: null output");
    *chan = v->clr_chan;
    return GVDBX_OK;
}

int gvdbx_add_channel(gvdbx_volume* v, int chan, int dtype, int apron, int filter, int border) {
    TRY(check_vol(v));
    if (!v->configured) FAIL(GVDBX_ERR_STATE, "AddChannel before Configure");
    if (chan < 0 || chan >= GVDBX_MAXCHAN) FAIL(GVDBX_ERR_ARG, "AddChannel: channel %d out of range", chan);
    if (dtype != GVDBX_FLOAT && dtype != GVDBX_UCHAR4)
        FAIL(GVDBX_ERR_ARG, "AddChannel: type %d is not supported (T_FLOAT = 3, T_UCHAR4 = 2)", dtype);
    if (apron < 0 || apron > 1) FAIL(GVDBX_ERR_ARG, "AddChannel: apron must be 0 or 1");
    if (!any_chan(v)) {
        for (int a = 0; a < 3; a++) v->subdim[a] = v->def_axis[a];
        v->apron = apron;
        v->atlas_cap = 0;
    } else if (apron != v->apron) {
        FAIL(GVDBX_ERR_ARG, "AddChannel: all channels share one apron");
    }
    int64_t slots = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
    int64_t cap = std::max(slots, v->atlas_cap);
    size_t bytes = (size_t)(cap * brick_vox(v)) * sizeof(float);
    free_buf(v->chan[chan]);
    v->buf_gen++;
    hipError_t e = hipMalloc(&v->chan[chan].p, bytes);
    if (e != hipSuccess) {
        v->chan[chan].p = nullptr;
        (void)hipGetLastError();
        FAIL(GVDBX_ERR_NOMEM, "AddChannel: hipMalloc(%zu) failed", bytes);
    }
    v->chan[chan].cap = bytes;
    chan_written(v, chan);
    v->chan_raw[chan] = false;
    launch_fill_f32(v->chan[chan].as<float>(), cap * brick_vox(v), 0.0f, v->stream);
    v->has_chan[chan] = true;
    v->chan_type[chan] = dtype;
    v->chan_filter[chan] = filter;
    v->chan_border[chan] = border;
    v->atlas_cap = cap;
    return check_launch("AddChannel");
}

int gvdbx_fill_channel(gvdbx_volume* v, int chan, float value) {
    TRY(check_vol(v));
    TRY(check_chan(v, chan));
    if (v->chan_type[chan] == GVDBX_UCHAR4) {
        // all zero clears (ClearChannel); otherwise gvdbOpFillC4
        // (cuda_gvdb_operators.cuh:368-373) stores (value, value, value, 1) * 255
        uint32_t w = 0u;
        if (value != 0.0f) {
            const uint32_t c = sat_u8(value * 255.0f);
            w = c | c << 8 | c << 16 | 255u << 24;
        }
        chan_written(v, chan);
        launch_fill_i32(v->chan[chan].as<int32_t>(), v->atlas_cap * brick_vox(v), (int32_t)w, v->stream);
        return check_launch("FillChannel");
    }
    chan_written(v, chan);
    launch_fill_f32(v->chan[chan].as<float>(), v->atlas_cap * brick_vox(v), value, v->stream);
    return check_launch("FillChannel");
}

int gvdbx_set_topology_mode(gvdbx_volume* v, int mode) {
    TRY(check_vol(v));
    if (mode < GVDBX_TOPO_AUTO || mode > GVDBX_TOPO_RADIX) FAIL(GVDBX_ERR_ARG, "unknown topology mode %d", mode);
    v->topo_mode = mode;
    return GVDBX_OK;
}

int gvdbx_set_tuning(gvdbx_volume* v, const char* key, int value) {
    TRY(check_vol(v));
    if (!key) FAIL(GVDBX_ERR_ARG, "null tuning key");
    if (!strcmp(key, "gather_variant")) {
        if (value < 0 || value > 2) FAIL(GVDBX_ERR_ARG, "gather_variant must be 0, 1 or 2");
        v->gather_variant = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "gather_split")) {
        if (value < 0 || value > (1 << 20)) FAIL(GVDBX_ERR_ARG, "gather_split must be in [0, 2^20]");
        v->gather_split = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "gather_classes")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "gather_classes must be 0 or 1");
        v->gather_classes = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "xcd_remap")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "xcd_remap must be 0 or 1");
        v->xcd_remap = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "splat_chunk")) {
        if (value < 64 || value > (1 << 24)) FAIL(GVDBX_ERR_ARG, "splat_chunk must be in [64, 2^24]");
        v->splat_chunk = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "splat_variant")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "splat_variant must be 0 or 1");
        v->splat_variant = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "sort_tiles")) {
        if (value < 0 || value > 2) FAIL(GVDBX_ERR_ARG, "sort_tiles must be 0, 1 or 2");
        v->sort_tiles = value;
        v->bins_valid = false;   // the next binning sorts again with the chosen tiles
        return GVDBX_OK;
    }
    if (!strcmp(key, "bin_records")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "bin_records must be 0 or 1");
        v->bin_records = value;
        v->bins_valid = false;
        return GVDBX_OK;
    }
    if (!strcmp(key, "apron_order")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "apron_order must be 0 or 1");
        v->apron_order = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "apron_push")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "apron_push must be 0 or 1");
        v->apron_push = value;
        return GVDBX_OK;
    }
    if (!strcmp(key, "gather_quarters")) {
        if (value < 0 || value > 1) FAIL(GVDBX_ERR_ARG, "gather_quarters must be 0 or 1");
        v->gather_quarters = value;
        return GVDBX_OK;
    }
    FAIL(GVDBX_ERR_ARG, "unknown tuning key '%s'", key);
}

int gvdbx_get_tuning(gvdbx_volume* v, const char* key, int* value) {
    TRY(check_vol(v));
    if (!key || !value) FAIL(GVDBX_ERR_ARG, "null tuning key or output");
    if (!strcmp(key, "gather_variant")) {
        *value = v->gather_variant;
        return GVDBX_OK;
    }
    if (!strcmp(key, "gather_split")) {
        *value = v->gather_split;
        return GVDBX_OK;
    }
    if (!strcmp(key, "gather_classes")) {
        *value = v->gather_classes;
        return GVDBX_OK;
    }
    if (!strcmp(key, "xcd_remap")) {
        *value = v->xcd_remap;
        return GVDBX_OK;
    }
    if (!strcmp(key, "splat_chunk")) {
        *value = v->splat_chunk;
        return GVDBX_OK;
    }
    if (!strcmp(key, "splat_variant")) {
        *value = v->splat_variant;
        return GVDBX_OK;
    }
    if (!strcmp(key, "sort_tiles")) {
        *value = v->sort_tiles;
        return GVDBX_OK;
    }
    if (!strcmp(key, "bin_records")) {
        *value = v->bin_records;
        return GVDBX_OK;
    }
    if (!strcmp(key, "apron_order")) {
        *value = v->apron_order;
        return GVDBX_OK;
    }
    if (!strcmp(key, "apron_push")) {
        *value = v->apron_push;
        return GVDBX_OK;
    }
    if (!strcmp(key, "gather_quarters")) {
        *value = v->gather_quarters;
        return GVDBX_OK;
    }
    FAIL(GVDBX_ERR_ARG, "unknown tuning key '%s'", key);
}

int gvdbx_set_points(gvdbx_volume* v, const void* pos_dev, int pos_off, int pos_stride, const void* vel_dev,
                     int vel_off, int vel_stride, const void* clr_dev, int clr_off, int clr_stride) {
    // velocities are accepted and not read: no stage of this path uses them
    // (the gathers copy them into the subcell lists but never read them,
    // cuda_gvdb_particles.cuh:696-773)
    (void)vel_dev; (void)vel_off; (void)vel_stride;
    TRY(check_vol(v));
    if (!pos_dev) FAIL(GVDBX_ERR_ARG, "SetPoints: null position pointer");
    if (pos_stride < 12 || pos_off < 0) FAIL(GVDBX_ERR_ARG, "SetPoints: position stride must be >= 12");
    if (((uintptr_t)pos_dev + pos_off) % 4 || pos_stride % 4) FAIL(GVDBX_ERR_ARG, "SetPoints: positions must be 4-byte aligned");
    if (clr_dev && (clr_off < 0 || clr_stride < 4 || clr_stride % 4 || ((uintptr_t)clr_dev + clr_off) % 4))
        FAIL(GVDBX_ERR_ARG, "SetPoints: colours must be 4-byte aligned uchar4 with a stride that is a multiple of 4");
    v->clr = clr_dev ? (const char*)clr_dev + clr_off : nullptr;
    v->clr_stride = clr_stride;
    v->pos.base = (const char*)pos_dev + pos_off;
    v->pos.stride = pos_stride;
    v->has_pos = true;
    v->pos_gen++;
    v->bins_valid = false;
    v->ncloud = 0;
    return GVDBX_OK;
}

int gvdbx_set_points_batched(gvdbx_volume* v, const void* pos_dev, int pos_stride, const int64_t* cloud_off,
                             int ncloud) {
    TRY(check_vol(v));
    if (!pos_dev || !cloud_off || ncloud < 1) FAIL(GVDBX_ERR_ARG, "SetPointsBatched: null pointer or no cloud");
    if (pos_stride < 12 || pos_stride % 4 || (uintptr_t)pos_dev % 4)
        FAIL(GVDBX_ERR_ARG, "SetPointsBatched: positions must be 4-byte aligned with stride >= 12");
    if (cloud_off[0] != 0) FAIL(GVDBX_ERR_ARG, "SetPointsBatched: cloud offsets must start at 0");
    for (int c = 0; c < ncloud; c++)
        if (cloud_off[c + 1] < cloud_off[c]) FAIL(GVDBX_ERR_ARG, "SetPointsBatched: cloud offsets must not decrease");
    if (cloud_off[ncloud] >= ((int64_t)1 << 31)) FAIL(GVDBX_ERR_RANGE, "SetPointsBatched: too many points");
    v->h_ptoff.assign(cloud_off, cloud_off + ncloud + 1);
    TRY(ensure(v, v->c_ptoff, (size_t)(ncloud + 1) * 8));
    TRY(ensure_host(v, (size_t)(ncloud + 1) * 8 + 4096));
    TRY(sync(v));
    memcpy(v->host, cloud_off, (size_t)(ncloud + 1) * 8);
    GVDBX_CHECK(hipMemcpyAsync(v->c_ptoff.p, v->host, (size_t)(ncloud + 1) * 8, hipMemcpyHostToDevice, v->stream));
    TRY(sync(v));
    v->pos.base = (const char*)pos_dev;
    v->pos.stride = pos_stride;
    v->clr = nullptr;
    v->has_pos = true;
    v->pos_gen++;
    v->bins_valid = false;
    v->ncloud = ncloud;
    v->h_cbounds.clear();
    return GVDBX_OK;
}

int gvdbx_read_cloud_layout(gvdbx_volume* v, int32_t* cell_off_host, float* bounds_host, int ncloud) {
    TRY(check_vol(v));
    if (v->ncloud == 0 || ncloud != v->ncloud) FAIL(GVDBX_ERR_ARG, "ReadCloudLayout: the volume holds %d clouds", v->ncloud);
    if (!v->has_topo) FAIL(GVDBX_ERR_STATE, "ReadCloudLayout: no topology");
    if (cell_off_host) memcpy(cell_off_host, v->h_celloff.data(), (size_t)ncloud * 12);
    if (bounds_host) {
        if ((int)v->h_cbounds.size() != 6 * ncloud) FAIL(GVDBX_ERR_STATE, "ReadCloudLayout: bounds not computed");
        memcpy(bounds_host, v->h_cbounds.data(), (size_t)ncloud * 24);
    }
    return GVDBX_OK;
}

int gvdbx_read_leaf_clouds(gvdbx_volume* v, int32_t* out_host, int64_t max_entries) {
    TRY(check_vol(v));
    if (v->ncloud == 0 || !v->has_topo) FAIL(GVDBX_ERR_STATE, "ReadLeafClouds: no batched topology");
    if (!out_host || max_entries < v->cnt[0]) FAIL(GVDBX_ERR_ARG, "ReadLeafClouds: output too small");
    GVDBX_CHECK(hipMemcpyAsync(out_host, v->c_leafcloud.p, (size_t)v->cnt[0] * 4, hipMemcpyDeviceToHost, v->stream));
    return sync(v);
}

int gvdbx_set_point_weights(gvdbx_volume* v, const float* w_dev, int stride) {
    TRY(check_vol(v));
    if (w_dev && (stride < 4 || stride % 4)) FAIL(GVDBX_ERR_ARG, "SetPointWeights: bad stride");
    v->wts = (const char*)w_dev;
    v->wstride = stride;
    return GVDBX_OK;
}

int gvdbx_rebuild_topology(gvdbx_volume* v, int num_pnts, float radius, float ox, float oy, float oz) {
    TRY(check_vol(v));
    float o[3] = {ox, oy, oz};
    // RequestFullRebuild(false): the incremental branch (VG.cpp:1195-1216)
    // while the depth holds and the points stay inside the root
    if (v->incre_rebuild && !v->force_full && v->has_topo && v->cnt[0] > 0 && v->ncloud == 0 &&
        v->top_lev == v->depth) {
        bool full = false;
        TRY(incremental_topology(v, num_pnts, radius, o, false, &full));
        if (!full) return GVDBX_OK;
    }
    return rebuild_topology(v, num_pnts, radius, o);
}

int gvdbx_accumulate_topology(gvdbx_volume* v, int num_pnts, float radius, float ox, float oy, float oz,
                              int min_depth) {
    TRY(check_vol(v));
    if (min_depth < 0) FAIL(GVDBX_ERR_ARG, "AccumulateTopology: negative minimum depth");
    float o[3] = {ox, oy, oz};
    // the first call, or one after RequestFullRebuild(true), rebuilds (VG.cpp:1173-1176)
    if (!v->has_topo || v->cnt[0] == 0 || v->force_full) return rebuild_topology(v, num_pnts, radius, o);
    return incremental_topology(v, num_pnts, radius, o, true, nullptr);
}

int gvdbx_request_full_rebuild(gvdbx_volume* v, int tf) {
    TRY(check_vol(v));
    v->incre_rebuild = !tf;
    v->force_full = tf != 0;
    return GVDBX_OK;
}

int gvdbx_finish_topology(gvdbx_volume* v, int commit_pool, int compute_bnds) {
    (void)commit_pool;   // pools live on the device only
    TRY(check_vol(v));
    if (!v->has_topo) FAIL(GVDBX_ERR_STATE, "FinishTopology: no topology");
    if (compute_bnds) TRY(compute_bounds(v));
    return GVDBX_OK;
}

int gvdbx_update_atlas(gvdbx_volume* v) {
    TRY(check_vol(v));
    return update_atlas(v);
}

}  // extern "C"

namespace {

// InsertPointsSubcell (VG.cpp:4971-5092) and InsertPointsSubcell_FP16
// (VG.cpp:5094-5189): the lists are the bins plus each point's subcell box,
// so only their length is computed here; fp16 records the box the FP16
// positions are fractions of (GetBoundingBox's pos_min and its range).
int insert_subcell(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty, float tz,
                   int64_t* out_len, bool fp16) {
    const char* what = fp16 ? "InsertPointsSubcell_FP16" : "InsertPointsSubcell";
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "%s: not available for batched clouds", what);
    TRY(need_atlas(v, what));
    if (num_pnts < 0) FAIL(GVDBX_ERR_ARG, "%s: negative point count", what);
    if (sc <= 0 || v->res[0] % sc) FAIL(GVDBX_ERR_ARG, "%s: subcell size must divide the brick", what);
    if (!(radius >= 0.0f) || radius + (float)sc > (float)v->res[0])
        FAIL(GVDBX_ERR_RANGE, "%s: radius %g + subcell %d exceeds the brick (%d)", what, radius, sc, v->res[0]);
    float t[3] = {tx, ty, tz};
    int64_t len = 0;
    TRY(ensure_bins(v, num_pnts, t, false));
    TRY(subcell_len(v, sc, radius, &len));
    v->sc_size = sc;
    v->sc_radius = radius;
    v->sc_len = len;
    v->sc_valid = true;
    v->sc_fp16 = fp16;
    v->sc_q = Quant16{};
    if (fp16) {
        v->sc_q.on = 1;
        for (int a = 0; a < 3; a++) {
            v->sc_q.mn[a] = v->pos_min[a];
            v->sc_q.rg[a] = v->pos_max[a] - v->pos_min[a];
        }
    }
    if (out_len) *out_len = len;
    return GVDBX_OK;
}

}  // namespace

extern "C" {

int gvdbx_insert_points_subcell(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty, float tz,
                                int64_t* out_len) {
    TRY(check_vol(v));
    return insert_subcell(v, sc, num_pnts, radius, tx, ty, tz, out_len, false);
}

int gvdbx_insert_points_subcell_fp16(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty,
                                     float tz, int64_t* out_len) {
    TRY(check_vol(v));
    return insert_subcell(v, sc, num_pnts, radius, tx, ty, tz, out_len, true);
}

int gvdbx_gather_level_set_fp16(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty, float tz,
                                int64_t sc_len, int chan_levelset, int chan_clr) {
    (void)num_pnts; (void)tx; (void)ty; (void)tz;
    TRY(check_vol(v));
    return gather(v, 1, sc, radius, sc_len, chan_levelset, chan_clr, 0, false, true);
}

int gvdbx_gather_density(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty, float tz,
                         int64_t sc_len, int chan_density, int chan_clr, int accum) {
    (void)num_pnts; (void)tx; (void)ty; (void)tz;   // the lists of InsertPointsSubcell fix the points
    TRY(check_vol(v));
    return gather(v, 0, sc, radius, sc_len, chan_density, chan_clr, accum);
}

int gvdbx_gather_level_set(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty, float tz,
                           int64_t sc_len, int chan_levelset, int chan_clr, int accum) {
    (void)num_pnts; (void)tx; (void)ty; (void)tz;
    TRY(check_vol(v));
    return gather(v, 1, sc, radius, sc_len, chan_levelset, chan_clr, accum);
}

int gvdbx_gather_level_set_noalloc(gvdbx_volume* v, int sc, int num_pnts, float radius, float tx, float ty, float tz,
                                   int64_t sc_len, int chan_levelset, int chan_clr, int accum) {
    (void)num_pnts; (void)tx; (void)ty; (void)tz;
    TRY(check_vol(v));
    return gather(v, 1, sc, radius, sc_len, chan_levelset, chan_clr, accum, true);
}

int gvdbx_insert_points(gvdbx_volume* v, int num_pnts, float tx, float ty, float tz, int prefix) {
    (void)prefix;
    TRY(check_vol(v));
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "InsertPoints: not available for batched clouds");
    TRY(need_atlas(v, "InsertPoints"));
    if (num_pnts < 0) FAIL(GVDBX_ERR_ARG, "InsertPoints: negative point count");
    float t[3] = {tx, ty, tz};
    TRY(ensure_bins(v, num_pnts, t, true));
    TRY(ensure(v, v->pnode, (size_t)std::max(num_pnts, 1) * 4));
    TRY(ensure(v, v->poff, (size_t)std::max(num_pnts, 1) * 4));
    launch_fill_i32(v->pnode.as<int32_t>(), num_pnts, (int32_t)GVDBX_ID_UNDEFL, v->stream);
    launch_fill_i32(v->poff.as<int32_t>(), num_pnts, 0, v->stream);
    InsertParams ip{v->gbin, v->grid[0], v->idx[0].as<int32_t>(), v->bin_off.as<int32_t>(), v->sorted.as<float4>(),
                    v->logdim[0], v->pnode.as<int32_t>(), v->poff.as<int32_t>(), num_pnts};
    launch_insert_points(ip, v->stream);
    v->npnode = num_pnts;
    return check_launch("InsertPoints");
}

int gvdbx_scatter_density(gvdbx_volume* v, int num_pnts, float radius, float amp, float tx, float ty, float tz,
                          int expand, int avg_color) {
    (void)amp;   // amp is unused by the reference kernel
    TRY(check_vol(v));
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "ScatterDensity: not available for batched clouds");
    TRY(need_atlas(v, "ScatterDensity"));
    TRY(check_chan_type(v, 0, GVDBX_FLOAT, "ScatterDensity"));
    // point colours go to channel 1 (averaged) or 3 (last point), the
    // channels gvdbScatterPointAvgCol / gvdbScatterPointDensity write
    // (cuda_gvdb_particles.cuh:570,641)
    const int cchan = avg_color ? 1 : 3;
    if (v->clr) TRY(check_chan_type(v, cchan, GVDBX_UCHAR4, "ScatterDensity"));
    if (num_pnts < 0) FAIL(GVDBX_ERR_ARG, "ScatterDensity: negative point count");
    float t[3] = {tx, ty, tz};
    TRY(ensure_bins(v, num_pnts, t, true));
    ScatterParams sp{v->nodes[0].as<GNode>(), v->cnt[0], v->gbin, v->bin_off.as<int32_t>(), v->sorted.as<float4>(),
                     v->chan[0].as<float>(), brick_res(v), std::max(v->apron, 0), v->res[0], radius, expand ? 1 : 0};
    if (sp.apron == 0 && expand) FAIL(GVDBX_ERR_ARG, "ScatterDensity: expand needs an apron");
    chan_written(v, 0);
    launch_scatter_density(sp, v->stream);
    if (v->clr) {
        ColorScatterParams cp{v->nodes[0].as<GNode>(), v->cnt[0], v->gbin, v->bin_off.as<int32_t>(),
                              v->sorted.as<float4>(), v->clr, v->clr_stride, v->chan[cchan].as<uint32_t>(),
                              brick_res(v), std::max(v->apron, 0), v->res[0], v->logdim[0], avg_color ? 1 : 0};
        chan_written(v, cchan);
        launch_scatter_color(cp, v->stream);
    }
    return check_launch("ScatterDensity");
}

// Cubic splat work runs over the current bins: each bin cell's points in
// runs of at most splat_chunk (sp_runoff = exclusive scan of the per-cell run
// counts, sp_runcell = run -> cell).  One host read of the run count.
static int splat_run_plan(gvdbx_volume* v, int* nruns_out) {
    hipStream_t s = v->stream;
    SplatParams sp{};
    sp.gb = v->gbin;
    sp.off = v->bin_off.as<int32_t>();
    const int64_t nc = v->gbin.count();
    const int64_t stiles = scan_tiles(nc + 1);
    TRY(ensure(v, v->sp_runs, (size_t)nc * 4));
    TRY(ensure(v, v->sp_runoff, (size_t)(nc + 1) * 4));
    TRY(ensure(v, v->sp_scan, (size_t)(stiles + 1) * 4));
    TRY(ensure_host(v, 4096));
    int32_t* runs = v->sp_runs.as<int32_t>();
    int32_t* run_off = v->sp_runoff.as<int32_t>();
    uint32_t* spart = v->sp_scan.as<uint32_t>();
    launch_splat_runs(sp, v->splat_chunk, runs, s);
    launch_count_scan_reduce(runs, nc, spart, spart + stiles, s);
    launch_count_scan_apply(runs, nc, spart, run_off, s);
    GVDBX_CHECK(hipMemcpyAsync(v->host, run_off + nc, 4, hipMemcpyDeviceToHost, s));
    TRY(sync(v));
    const int nruns = *(const int32_t*)v->host;
    TRY(ensure(v, v->sp_runcell, (size_t)std::max(nruns, 1) * 4));
    launch_splat_run_map(sp, run_off, v->sp_runcell.as<int32_t>(), s);
    *nruns_out = nruns;
    return GVDBX_OK;
}

int gvdbx_splat_cubic(gvdbx_volume* v, int num_pnts, float radius, float tx, float ty, float tz, int chan, int accum) {
    TRY(check_vol(v));
    TRY(need_atlas(v, "SplatCubic"));
    TRY(check_chan_type(v, chan, GVDBX_FLOAT, "SplatCubic"));
    if (num_pnts < 0) FAIL(GVDBX_ERR_ARG, "SplatCubic: negative point count");
    if (!(radius > 0.0f) || radius > (float)v->res[0]) FAIL(GVDBX_ERR_RANGE, "SplatCubic: radius must be in (0, brick]");
    float t[3] = {tx, ty, tz};
    TRY(ensure_bins(v, num_pnts, t, true));
    SplatParams sp{};
    if (v->ncloud > 0) {
        sp.leaf_cloud = v->c_leafcloud.as<int32_t>();
        sp.cell_off = v->c_celloff.as<int32_t>();
    }
    sp.leaves = v->nodes[0].as<GNode>();
    sp.nleaf = v->cnt[0];
    sp.gb = v->gbin;
    sp.off = v->bin_off.as<int32_t>();
    sp.weights = v->wts;
    sp.wstride = v->wstride;
    sp.chan = v->chan[chan].as<float>();
    chan_written(v, chan);
    sp.brick_res = brick_res(v);
    sp.apron = std::max(v->apron, 0);
    sp.accum = accum ? 1 : 0;
    sp.radius = radius;
    sp.bins = v->sorted.as<float4>();
    sp.cm = cloud_map(v);
    hipStream_t s = v->stream;
    const int res0 = v->res[0];
    if (v->splat_variant == 0 || splat_cell_lds(res0, radius) > 65536) {
        TRY(ensure(v, v->splat_src, (size_t)std::max(num_pnts, 1) * 16));
        launch_splat_src(v->sorted.as<float4>(), num_pnts, v->bin_off.as<int32_t>() + v->gbin.count(), v->wts,
                         v->wstride, v->splat_src.as<float4>(), s);
        sp.sorted = v->splat_src.as<float4>();
        launch_splat_tile(sp, s);
        return check_launch("SplatCubic");
    }
    int nruns = 0;
    TRY(splat_run_plan(v, &nruns));
    const int64_t G = res0 + 2 * splat_halo(radius);
    const int32_t* run_off = v->sp_runoff.as<int32_t>();
    TRY(ensure(v, v->sp_part, (size_t)std::max(nruns, 1) * G * G * G * 4));
    launch_splat_cell(sp, v->sp_runcell.as<int32_t>(), run_off, nruns, v->splat_chunk, v->sp_part.as<float>(), s);
    launch_splat_collect(sp, run_off, v->sp_part.as<float>(), s);
    return check_launch("SplatCubic");
}

}  // extern "C"

namespace {

int update_apron(gvdbx_volume* v, int chan, float boundval) {
    TRY(check_chan(v, chan));
    if (v->apron == 0) return GVDBX_OK;   // VG.cpp:4420
    TRY(need_atlas(v, "UpdateApron"));
    // the word stored outside every leaf: the float itself, or
    // make_uchar4(b, b, b, b) for colour channels (cuda_gvdb_operators.cuh:139-142)
    uint32_t bits;
    if (v->chan_type[chan] == GVDBX_UCHAR4) {
        const uint32_t c = sat_u8(boundval);
        bits = c | c << 8 | c << 16 | c << 24;
    } else {
        memcpy(&bits, &boundval, 4);
    }
    const bool pushed = v->apron_pushed[chan] && v->apron_push_gen[chan] == v->topo_gen && !v->chan_raw[chan];
    ApronParams ap{v->nodes[0].as<GNode>(), v->cnt[0], v->grid[0], v->idx[0].as<int32_t>(), v->res[0],
                   v->logdim[0], v->chan[chan].as<float>(), brick_res(v), boundval, bits, v->apron_order,
                   pushed ? 1 : 0};
    launch_apron(ap, v->stream);
    return check_launch("UpdateApron");
}

}  // namespace

extern "C" {

int gvdbx_update_apron(gvdbx_volume* v, int chan, float boundval) {
    TRY(check_vol(v));
    return update_apron(v, chan, boundval);
}

int gvdbx_render(gvdbx_volume* v, int shade, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp,
                 float* rgba_dev, uint8_t* rgba8_dev) {
    TRY(check_vol(v));
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "Render: not available for batched clouds");
    if (!(shade >= GVDBX_SHADE_VOXEL && shade <= GVDBX_SHADE_VOLUME) && shade != GVDBX_SHADE_OFF)
        FAIL(GVDBX_ERR_ARG, "Render: unknown shading %d", shade);
    RenderTree tv;
    RenderCam rc;
    RenderScene sc;
    if (shade == GVDBX_SHADE_OFF) {   // VG.cpp:4340-4345: cleared before anything is prepared
        if (!cam) FAIL(GVDBX_ERR_ARG, "Render: null camera");
        if (cam->width <= 0 || cam->height <= 0 || (int64_t)cam->width * cam->height >= ((int64_t)1 << 31))
            FAIL(GVDBX_ERR_ARG, "Render: bad image size %dx%d", cam->width, cam->height);
        memset(&tv, 0, sizeof(tv));
        memset(&rc, 0, sizeof(rc));
        memset(&sc, 0, sizeof(sc));
        rc.width = cam->width;
        rc.height = cam->height;
    } else {
        TRY(make_render(v, chan, cam, rp, "Render", tv, rc, sc));
        const bool lut = shade == GVDBX_SHADE_VOLUME || shade == GVDBX_SHADE_SECTION2D || shade == GVDBX_SHADE_SECTION3D;
        if (lut && !rp->lut_dev) FAIL(GVDBX_ERR_ARG, "Render: shading %d needs a transfer function", shade);
        if (((uintptr_t)rp->lut_dev & 15) != 0) FAIL(GVDBX_ERR_ARG, "Render: transfer function must be 16-byte aligned");
        if (shade == GVDBX_SHADE_TRICUBIC && !(rp->finestep > 0.0f && isfinite(rp->finestep)))
            FAIL(GVDBX_ERR_ARG, "Render: SHADE_TRICUBIC needs finestep > 0");
        if (v->clr_chan >= 0) {
            TRY(check_chan_type(v, v->clr_chan, GVDBX_UCHAR4, "Render (colour channel)"));
            sc.clr = v->chan[v->clr_chan].as<uchar4>();
        }
    }
    if (!rgba_dev && !rgba8_dev) return GVDBX_OK;
    launch_render(shade, tv, rc, sc, (float4*)rgba_dev, (uchar4*)rgba8_dev, v->stream);
    return check_launch("Render");
}

int gvdbx_raytrace(gvdbx_volume* v, int chan, const gvdbx_render_params* rp, gvdbx_scn_ray* rays_dev, int64_t num_rays,
                   float bias) {
    TRY(check_vol(v));
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "Raytrace: not available for batched clouds");
    if (num_rays < 0 || (num_rays > 0 && !rays_dev)) FAIL(GVDBX_ERR_ARG, "Raytrace: bad ray buffer");
    if ((uintptr_t)rays_dev & 15) FAIL(GVDBX_ERR_ARG, "Raytrace: ray records must be 16-byte aligned");
    gvdbx_camera one;
    memset(&one, 0, sizeof(one));
    one.width = one.height = 1;   // PrepareRender(1, 1, 0) VG.cpp:4390
    RenderTree tv;
    RenderCam rc;
    RenderScene sc;
    TRY(make_render(v, chan, &one, rp, "Raytrace", tv, rc, sc));
    launch_raytrace(tv, sc, (GScnRay*)rays_dev, num_rays, bias, v->stream);
    return check_launch("Raytrace");
}

}  // extern "C"

namespace {

// Upload one camera per cloud, each with its cloud's shift and bounds.
int upload_cams(gvdbx_volume* v, const gvdbx_camera* cams, int ncam, const char* what) {
    if (v->ncloud == 0) FAIL(GVDBX_ERR_STATE, "%s: the volume holds no batched clouds", what);
    if (ncam != v->ncloud) FAIL(GVDBX_ERR_ARG, "%s: %d cameras for %d clouds", what, ncam, v->ncloud);
    if ((int)v->h_cbounds.size() != 6 * ncam) FAIL(GVDBX_ERR_STATE, "%s: FinishTopology(compute_bounds) has not run", what);
    std::vector<RenderCam> rc((size_t)ncam);
    for (int c = 0; c < ncam; c++) {
        if (cams[c].width != cams[0].width || cams[c].height != cams[0].height)
            FAIL(GVDBX_ERR_ARG, "%s: every camera must have the same image size", what);
        fill_cam(&cams[c], rc[c]);
        for (int a = 0; a < 3; a++) {
            rc[c].shift[a] = v->h_celloff[3 * c + a] * v->res[0];
            rc[c].bmin[a] = v->h_cbounds[6 * c + a];
            rc[c].bmax[a] = v->h_cbounds[6 * c + 3 + a];
        }
        rc[c].has_bounds = 1;
    }
    TRY(ensure(v, v->c_cams, (size_t)ncam * sizeof(RenderCam)));
    TRY(ensure_host(v, (size_t)ncam * sizeof(RenderCam) + 4096));
    TRY(sync(v));   // the staging buffer may still feed an earlier copy
    memcpy(v->host, rc.data(), (size_t)ncam * sizeof(RenderCam));
    GVDBX_CHECK(hipMemcpyAsync(v->c_cams.p, v->host, (size_t)ncam * sizeof(RenderCam), hipMemcpyHostToDevice, v->stream));
    return GVDBX_OK;
}

int raycast_ea(gvdbx_volume* v, int chan, const gvdbx_camera* cams, int ncam, bool batched,
               const gvdbx_render_params* rp, float* I_dev, float* T_dev, int record) {
    TRY(check_vol(v));
    const char* what = batched ? "RaycastEABatched" : "RaycastEA";
    if (!cams || ncam < 1) FAIL(GVDBX_ERR_ARG, "%s: no camera", what);
    RenderTree tv;
    RenderCam rc;
    RenderScene sc;
    TRY(make_render(v, chan, &cams[0], rp, what, tv, rc, sc));
    if (!I_dev || !T_dev) FAIL(GVDBX_ERR_ARG, "%s: null output", what);
    if (!batched && v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "%s: the volume holds batched clouds (one camera each)", what);
    const RenderCam* dcams = nullptr;
    if (batched) {
        TRY(upload_cams(v, cams, ncam, what));
        dcams = v->c_cams.as<RenderCam>();
    }
    hipStream_t s = v->stream;
    VisitRec vr{};
    if (!record) {
        launch_raycast_ea(tv, rc, dcams, ncam, sc, I_dev, T_dev, 0, vr, s);
        return check_launch(what);
    }
    const int64_t rays = (int64_t)rc.width * rc.height * (batched ? ncam : 1);
    if (rays >= ((int64_t)1 << 32)) FAIL(GVDBX_ERR_RANGE, "%s: too many rays", what);
    const int lb = bits_for(v->cnt[0]);
    // one checkpoint chunk per flush of a wave's buffer: every chunk but a
    // wave's last holds more than 64 records
    const int64_t nwaves = (int64_t)((rc.width + 15) / 16) * ((rc.height + 15) / 16) * (batched ? ncam : 1) * 4;
    if (nwaves >= ((int64_t)1 << 32)) FAIL(GVDBX_ERR_RANGE, "%s: too many rays", what);
    v->vis_valid = false;
    int64_t cap = std::max<int64_t>(v->vis_cap, std::max<int64_t>(rays * 8, 1 << 20));
    // vis_cnt: record and chunk counters (8 B), padding, then one record
    // total per wave
    TRY(ensure(v, v->vis_cnt, (size_t)(4 + nwaves) * 4));
    TRY(ensure_host(v, 4096));
    for (int attempt = 0; attempt < 2; attempt++) {
        if (cap >= ((int64_t)1 << 31)) FAIL(GVDBX_ERR_RANGE, "%s: more than 2^31 segment checkpoints", what);
        const int64_t ccap = cap / 64 + nwaves + 16;
        TRY(ensure(v, v->vis_key, (size_t)cap * 4));
        TRY(ensure(v, v->vis_rec, (size_t)cap * 16));
        TRY(ensure(v, v->vis_ctag, (size_t)ccap * 8));
        TRY(ensure(v, v->vis_cbase, (size_t)ccap * 4));
        TRY(ensure(v, v->vis_csize, (size_t)ccap * 4));
        launch_fill_i32(v->vis_cnt.as<int32_t>(), 4 + nwaves, 0, s);
        vr.key = v->vis_key.as<uint32_t>();
        vr.rec = v->vis_rec.as<float4>();
        vr.count = v->vis_cnt.as<uint32_t>();
        vr.cap = cap;
        vr.chunk_tag = v->vis_ctag.as<uint64_t>();
        vr.chunk_base = v->vis_cbase.as<uint32_t>();
        vr.chunk_size = v->vis_csize.as<uint32_t>();
        vr.chunk_cap = ccap;
        vr.wave_recs = v->vis_cnt.as<uint32_t>() + 4;
        launch_raycast_ea(tv, rc, dcams, ncam, sc, I_dev, T_dev, 1, vr, s);
        TRY(check_launch(what));
        GVDBX_CHECK(hipMemcpyAsync(v->host, v->vis_cnt.p, 8, hipMemcpyDeviceToHost, s));
        TRY(sync(v));
        const uint32_t n = ((const uint32_t*)v->host)[0], m = ((const uint32_t*)v->host)[1];
        if ((int64_t)n <= cap && (int64_t)m <= ccap) {
            v->vis_n = n;
            v->vis_nchunk = m;
            v->vis_cap = cap;
            v->vis_ccap = ccap;
            v->vis_leaf_bits = lb;
            v->vis_nwaves = nwaves;
            v->vis_w = rc.width;
            v->vis_h = rc.height;
            v->vis_ncam = batched ? ncam : 0;
            v->vis_chan = chan;
            v->vis_topo_gen = v->topo_gen;
            v->vis_valid = true;
            v->vis_sorted = false;
            return GVDBX_OK;
        }
        cap = std::max<int64_t>(cap, (int64_t)n + (int64_t)n / 8 + 1024);   // the segment count of this view; rerun once
    }
    FAIL(GVDBX_ERR_STATE, "%s: segment count did not settle", what);
}

int raycast_ea_backward(gvdbx_volume* v, int chan, const gvdbx_camera* cams, int ncam, bool batched,
                        const gvdbx_render_params* rp, const float* I_dev, const float* T_dev, const float* gI_dev,
                        const float* gT_dev, float* grad_dev) {
    TRY(check_vol(v));
    const char* what = batched ? "RaycastEABackwardBatched" : "RaycastEABackward";
    if (!cams || ncam < 1) FAIL(GVDBX_ERR_ARG, "%s: no camera", what);
    RenderTree tv;
    RenderCam rc;
    RenderScene sc;
    TRY(make_render(v, chan, &cams[0], rp, what, tv, rc, sc));
    if (!I_dev || !T_dev || !gI_dev || !grad_dev) FAIL(GVDBX_ERR_ARG, "%s: null buffer", what);
    if (!v->vis_valid || v->vis_topo_gen != v->topo_gen || v->vis_w != rc.width || v->vis_h != rc.height ||
        v->vis_chan != chan || v->vis_ncam != (batched ? ncam : 0))
        FAIL(GVDBX_ERR_STATE, "%s: no recorded raycast for this view and topology", what);
    if (brick_res(v) != 10) FAIL(GVDBX_ERR_ARG, "%s: needs 8^3 bricks with apron 1", what);
    const RenderCam* dcams = nullptr;
    if (batched) {
        TRY(upload_cams(v, cams, ncam, what));
        dcams = v->c_cams.as<RenderCam>();
    }
    hipStream_t s = v->stream;
    const int64_t n = v->vis_n, nleaf = v->cnt[0];
    if (!v->vis_sorted) {
        // Checkpoints grouped by leaf in an order fixed by the data: records
        // laid out in (wave, march) order (each chunk placed at its wave's
        // start, from a scan of the waves' record totals, plus the records
        // its wave wrote before it), then a stable sort by leaf.  Within a
        // leaf the segments are then in (wave, march) order whichever wave
        // wrote first.
        const int64_t m = v->vis_nchunk, nw = v->vis_nwaves;
        const int64_t nh = radix_hist_entries(n);
        TRY(ensure(v, v->vis_coff, (size_t)(nw + 1) * 4));
        TRY(ensure(v, v->vis_lkey, (size_t)std::max<int64_t>(n, 1) * 4));
        TRY(ensure(v, v->vis_idx, (size_t)std::max<int64_t>(n, 1) * 4));
        TRY(ensure(v, v->vis_tkey, (size_t)std::max<int64_t>(n, 1) * 4));
        TRY(ensure(v, v->vis_tidx, (size_t)std::max<int64_t>(n, 1) * 4));
        TRY(ensure(v, v->vis_hist, (size_t)nh * 4));
        TRY(ensure(v, v->vis_off, (size_t)(nh + 1) * 4));
        TRY(ensure(v, v->vis_part, (size_t)(scan_tiles(std::max(nh, nw) + 1) + 1) * 4));
        const int32_t* wrec = v->vis_cnt.as<int32_t>() + 4;
        int32_t* woff = v->vis_coff.as<int32_t>();
        uint32_t* spart = v->vis_part.as<uint32_t>();
        launch_count_scan_reduce(wrec, nw, spart, spart + scan_tiles(nw + 1), s);
        launch_count_scan_apply(wrec, nw, spart, woff, s);
        launch_chunk_order(v->vis_ctag.as<uint64_t>(), m, v->vis_cbase.as<uint32_t>(), v->vis_csize.as<uint32_t>(),
                           woff, v->vis_key.as<uint32_t>(), v->vis_lkey.as<uint32_t>(), v->vis_idx.as<uint32_t>(), s);
        const int flip = radix_sort_pairs32(v->vis_lkey.as<uint32_t>(), v->vis_idx.as<uint32_t>(),
                                            v->vis_tkey.as<uint32_t>(), v->vis_tidx.as<uint32_t>(), n, 0,
                                            v->vis_leaf_bits, v->vis_hist.as<int32_t>(), v->vis_off.as<int32_t>(),
                                            v->vis_part.as<uint32_t>(), s);
        TRY(check_launch("segment order"));
        v->vis_in_tmp = flip != 0;
        TRY(ensure(v, v->vis_seg, (size_t)std::max<int64_t>(2 * nleaf, 2) * 4));
        launch_fill_i32(v->vis_seg.as<int32_t>(), 2 * nleaf, 0, s);
        const uint32_t* skey = v->vis_in_tmp ? v->vis_tkey.as<uint32_t>() : v->vis_lkey.as<uint32_t>();
        launch_segments(skey, n, v->vis_seg.as<int32_t>(), nleaf, s);
        TRY(check_launch("segment ranges"));
        v->vis_sorted = true;
    }
    EABackParams p{};
    p.tv = tv;
    p.cam = rc;
    p.cams = dcams;
    p.sc = sc;
    p.idx = v->vis_in_tmp ? v->vis_tidx.as<uint32_t>() : v->vis_idx.as<uint32_t>();
    p.rec = v->vis_rec.as<float4>();
    p.seg = v->vis_seg.as<int32_t>();
    p.I = I_dev;
    p.T = T_dev;
    p.gI = gI_dev;
    p.gT = gT_dev;
    p.grad = grad_dev;
    launch_ea_backward(p, s);
    return check_launch(what);
}

}  // namespace

extern "C" {

int gvdbx_raycast_ea(gvdbx_volume* v, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp,
                     float* I_dev, float* T_dev, int record) {
    return raycast_ea(v, chan, cam, 1, false, rp, I_dev, T_dev, record);
}

int gvdbx_raycast_ea_backward(gvdbx_volume* v, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp,
                              const float* I_dev, const float* T_dev, const float* gI_dev, const float* gT_dev,
                              float* grad_dev) {
    return raycast_ea_backward(v, chan, cam, 1, false, rp, I_dev, T_dev, gI_dev, gT_dev, grad_dev);
}

int gvdbx_raycast_ea_batched(gvdbx_volume* v, int chan, const gvdbx_camera* cams, int ncam,
                             const gvdbx_render_params* rp, float* I_dev, float* T_dev, int record) {
    return raycast_ea(v, chan, cams, ncam, true, rp, I_dev, T_dev, record);
}

int gvdbx_raycast_ea_backward_batched(gvdbx_volume* v, int chan, const gvdbx_camera* cams, int ncam,
                                      const gvdbx_render_params* rp, const float* I_dev, const float* T_dev,
                                      const float* gI_dev, const float* gT_dev, float* grad_dev) {
    return raycast_ea_backward(v, chan, cams, ncam, true, rp, I_dev, T_dev, gI_dev, gT_dev, grad_dev);
}

int gvdbx_apron_adjoint(gvdbx_volume* v, float* grad_dev) {
    TRY(check_vol(v));
    TRY(need_atlas(v, "ApronAdjoint"));
    if (!grad_dev) FAIL(GVDBX_ERR_ARG, "ApronAdjoint: null gradient");
    if (v->apron != 1) FAIL(GVDBX_ERR_ARG, "ApronAdjoint: needs apron 1");
    ApronParams ap{v->nodes[0].as<GNode>(), v->cnt[0], v->grid[0], v->idx[0].as<int32_t>(), v->res[0],
                   v->logdim[0], nullptr, brick_res(v), 0.0f};
    launch_apron_adjoint(ap, grad_dev, v->stream);
    return check_launch("ApronAdjoint");
}

int gvdbx_splat_cubic_backward(gvdbx_volume* v, int num_pnts, float radius, float tx, float ty, float tz,
                               const float* grad_dev, float* dpos_dev, float* dw_dev) {
    TRY(check_vol(v));
    TRY(need_atlas(v, "SplatCubicBackward"));
    if (num_pnts < 0) FAIL(GVDBX_ERR_ARG, "SplatCubicBackward: negative point count");
    if (!grad_dev || (!dpos_dev && num_pnts > 0)) FAIL(GVDBX_ERR_ARG, "SplatCubicBackward: null buffer");
    if (!(radius > 0.0f) || radius > (float)v->res[0])
        FAIL(GVDBX_ERR_RANGE, "SplatCubicBackward: radius must be in (0, brick]");
    if (v->apron < 1) FAIL(GVDBX_ERR_ARG, "SplatCubicBackward: needs an apron");
    if (num_pnts == 0) return GVDBX_OK;
    float t[3] = {tx, ty, tz};
    TRY(ensure_bins(v, num_pnts, t, false));
    hipStream_t s = v->stream;
    launch_fill_f32(dpos_dev, 3 * (int64_t)num_pnts, 0.0f, s);
    if (dw_dev) launch_fill_f32(dw_dev, num_pnts, 0.0f, s);
    SplatBackParams p{};
    p.cm = cloud_map(v);
    p.g0 = v->grid[0];
    p.idx0 = v->idx[0].as<int32_t>();
    p.leaves = v->nodes[0].as<GNode>();
    p.nleaf = v->cnt[0];
    p.logr0 = v->logdim[0];
    p.brick_res = brick_res(v);
    p.sorted = v->sorted.as<float4>();
    p.total = v->bin_off.as<int32_t>() + v->gbin.count();
    p.nmax = num_pnts;
    p.weights = v->wts;
    p.wstride = v->wstride;
    p.radius = radius;
    p.grad = grad_dev;
    p.dpos = dpos_dev;
    p.dw = dw_dev;
    const int res0 = v->res[0];
    if (v->splat_variant == 0 || splat_cell_lds(res0, radius) > 65536) {
        launch_splat_backward(p, s);
        return check_launch("SplatCubicBackward");
    }
    int nruns = 0;
    TRY(splat_run_plan(v, &nruns));
    launch_splat_cell_backward(p, v->gbin, v->bin_off.as<int32_t>(), v->sp_runcell.as<int32_t>(),
                               v->sp_runoff.as<int32_t>(), nruns, res0, v->splat_chunk, s);
    return check_launch("SplatCubicBackward");
}

int gvdbx_backproject_depth(const float* depth_dev, int nframes, int width, int height, const float* intr4,
                            const float* poses_dev, float voxel, const float* origin3, float dmin, float dmax,
                            float* out_xyz_dev, void* stream) {
    if (nframes < 0 || width < 0 || height < 0) FAIL(GVDBX_ERR_ARG, "BackprojectDepth: negative size");
    if ((int64_t)nframes * width * height == 0) return GVDBX_OK;
    if (!depth_dev || !intr4 || !poses_dev || !origin3 || !out_xyz_dev) FAIL(GVDBX_ERR_ARG, "BackprojectDepth: null pointer");
    if (!(voxel > 0.0f) || !isfinite(voxel)) FAIL(GVDBX_ERR_ARG, "BackprojectDepth: voxel size must be > 0");
    if (!(intr4[0] != 0.0f) || !(intr4[1] != 0.0f)) FAIL(GVDBX_ERR_ARG, "BackprojectDepth: zero focal length");
    launch_backproject(depth_dev, nframes, width, height, intr4, poses_dev, voxel, origin3, dmin, dmax, out_xyz_dev,
                       (hipStream_t)stream);
    return check_launch("BackprojectDepth");
}

int gvdbx_get_info(gvdbx_volume* v, gvdbx_info* out) {
    TRY(check_vol(v));
    if (!out) FAIL(GVDBX_ERR_ARG, "null info pointer");
    memset(out, 0, sizeof(*out));
    out->levs = v->levs;
    out->depth = v->depth;
    out->top_lev = v->top_lev;
    memcpy(out->root_pos, v->root_pos, sizeof(out->root_pos));
    for (int l = 0; l < GVDBX_MAXLEV; l++) out->node_cnt[l] = v->cnt[l];
    memcpy(out->subdim, v->subdim, sizeof(out->subdim));
    out->brick_res = brick_res(v);
    out->apron = std::max(v->apron, 0);
    memcpy(out->pos_min, v->pos_min, sizeof(out->pos_min));
    memcpy(out->pos_max, v->pos_max, sizeof(out->pos_max));
    TRY(resolve_bounds(v));
    memcpy(out->obj_min, v->obj_min, sizeof(out->obj_min));
    memcpy(out->obj_max, v->obj_max, sizeof(out->obj_max));
    out->sc_len = v->sc_len;
    out->topology_mode_used = v->topo_mode_used;
    for (int l = 0; l < GVDBX_MAXLEV && l < 10; l++) out->log2dim[l] = v->logdim[l];
    return GVDBX_OK;
}

int gvdbx_read_nodes(gvdbx_volume* v, int lev, void* out_host, int64_t max_nodes) {
    TRY(check_vol(v));
    if (!v->has_topo) FAIL(GVDBX_ERR_STATE, "no topology");
    if (lev < 0 || lev >= v->levs || !out_host) FAIL(GVDBX_ERR_ARG, "bad level or buffer");
    int64_t n = v->cnt[lev];
    if (max_nodes < n) FAIL(GVDBX_ERR_ARG, "buffer holds %lld nodes, need %lld", (long long)max_nodes, (long long)n);
    if (n == 0) return GVDBX_OK;
    GVDBX_CHECK(hipMemcpyAsync(out_host, v->nodes[lev].p, (size_t)n * sizeof(GNode), hipMemcpyDeviceToHost, v->stream));
    return sync(v);
}

int gvdbx_read_child_lists(gvdbx_volume* v, int lev, uint64_t* out_host, int64_t max_entries) {
    TRY(check_vol(v));
    if (!v->has_topo) FAIL(GVDBX_ERR_STATE, "no topology");
    if (lev < 1 || lev >= v->levs || !out_host) FAIL(GVDBX_ERR_ARG, "bad level or buffer");
    int64_t len = (int64_t)v->res[lev] * v->res[lev] * v->res[lev];
    int64_t n = v->cnt[lev] * len;
    if (max_entries < n) FAIL(GVDBX_ERR_ARG, "buffer too small");
    if (n == 0) return GVDBX_OK;
    GVDBX_CHECK(hipMemcpyAsync(out_host, v->clist[lev].p, (size_t)n * 8, hipMemcpyDeviceToHost, v->stream));
    return sync(v);
}

int gvdbx_read_atlas_map(gvdbx_volume* v, void* out_host, int64_t max_entries) {
    TRY(check_vol(v));
    if (!v->has_atlas) FAIL(GVDBX_ERR_STATE, "no atlas map (UpdateAtlas)");
    if (!out_host || max_entries < v->amap_cnt) FAIL(GVDBX_ERR_ARG, "buffer too small");
    GVDBX_CHECK(hipMemcpyAsync(out_host, v->amap.p, (size_t)v->amap_cnt * sizeof(GAtlasNode), hipMemcpyDeviceToHost,
                               v->stream));
    return sync(v);
}

int gvdbx_read_bricks(gvdbx_volume* v, int chan, float* out_host, int64_t max_floats) {
    TRY(check_vol(v));
    TRY(check_chan(v, chan));
    if (!v->has_atlas) FAIL(GVDBX_ERR_STATE, "no atlas (UpdateAtlas)");
    int64_t n = v->cnt[0] * brick_vox(v);
    if (!out_host || max_floats < n) FAIL(GVDBX_ERR_ARG, "buffer too small");
    if (n == 0) return GVDBX_OK;
    GVDBX_CHECK(hipMemcpyAsync(out_host, v->chan[chan].p, (size_t)n * 4, hipMemcpyDeviceToHost, v->stream));
    return sync(v);
}

int gvdbx_channel_device_ptr(gvdbx_volume* v, int chan, float** out_dev) {
    TRY(check_vol(v));
    TRY(check_chan(v, chan));
    if (!out_dev) FAIL(GVDBX_ERR_ARG, "null output");
    *out_dev = v->chan[chan].as<float>();
    v->chan_raw[chan] = true;   // the caller may write any voxel from now on
    chan_written(v, chan);
    return GVDBX_OK;
}

int gvdbx_read_point_nodes(gvdbx_volume* v, int32_t* pnode_host, int32_t* poff_host, int64_t n) {
    TRY(check_vol(v));
    if (n > v->npnode) FAIL(GVDBX_ERR_ARG, "only %lld points were inserted", (long long)v->npnode);
    if (n == 0) return GVDBX_OK;
    if (pnode_host)
        GVDBX_CHECK(hipMemcpyAsync(pnode_host, v->pnode.p, (size_t)n * 4, hipMemcpyDeviceToHost, v->stream));
    if (poff_host) GVDBX_CHECK(hipMemcpyAsync(poff_host, v->poff.p, (size_t)n * 4, hipMemcpyDeviceToHost, v->stream));
    return sync(v);
}

}  // extern "C"

// ---------------- VBX files (SaveVBX VG.cpp:1626-1767, LoadVBX 507-685) ----------------
namespace {

struct VbxIO {
    FILE* fp = nullptr;
    bool ok = true;
    void w(const void* p, size_t n) {
        if (ok && n && fwrite(p, 1, n, fp) != n) ok = false;
    }
    bool r(void* p, size_t n) {
        if (ok && n && fread(p, 1, n, fp) != n) ok = false;
        return ok;
    }
    ~VbxIO() {
        if (fp) fclose(fp);
    }
};

// channels in use must be 0 .. n-1 (the file numbers them by position)
int vbx_channels(gvdbx_volume* v, const char* what, int* out) {
    int nch = 0;
    while (nch < GVDBX_MAXCHAN && v->has_chan[nch]) nch++;
    for (int c = nch; c < GVDBX_MAXCHAN; c++)
        if (v->has_chan[c]) FAIL(GVDBX_ERR_ARG, "%s: channels must be numbered 0..n-1", what);
    *out = nch;
    return GVDBX_OK;
}

int dev_to_host(gvdbx_volume* v, void* dst, const void* src, size_t bytes) {
    if (!bytes) return GVDBX_OK;
    GVDBX_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, v->stream));
    return sync(v);
}

int host_to_dev(gvdbx_volume* v, void* dst, const void* src, size_t bytes) {
    if (!bytes) return GVDBX_OK;
    GVDBX_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, v->stream));
    return sync(v);
}

// One grid of LoadVBX: Configure, the pools, FinishTopology, the channels
// and UpdateAtlas (VG.cpp:596-679).  The atlas is taken as it lies in the
// file: each active leaf gets the brick at the slot UpdateAtlas gives it.
int vbx_load_grid(gvdbx_volume* v, VbxIO& f, const char* path) {
    char name[256];
    unsigned char dtype = 0, comps = 0, compress = 0, topotype = 0, layout = 0;
    float vsize[3];
    int leafcnt = 0, leafdim[3], apron = 0, nch = 0, reuse = 0, axiscnt[3], axisres[3], levels = 0;
    uint64_t atlas_sz = 0, root = 0;
    f.r(name, 256), f.r(&dtype, 1), f.r(&comps, 1), f.r(&compress, 1), f.r(vsize, 12), f.r(&leafcnt, 4);
    f.r(leafdim, 12), f.r(&apron, 4), f.r(&nch, 4), f.r(&atlas_sz, 8), f.r(&topotype, 1), f.r(&reuse, 4);
    f.r(&layout, 1), f.r(axiscnt, 12), f.r(axisres, 12), f.r(&levels, 4), f.r(&root, 8);
    if (!f.ok) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated grid header", path);
    if (dtype != 'f' || comps != 1 || compress != 0 || topotype != 2 || layout != 0)
        FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: only uncompressed float grids with GVDB topology in atlas layout", path);
    if (levels < 2 || levels > GVDBX_MAXLEV) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: %d levels", path, levels);
    if (nch < 0 || nch > GVDBX_MAXCHAN || apron < 0 || apron > 1)
        FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: %d channels with apron %d", path, nch, apron);
    int ld[GVDBX_MAXLEV], res[GVDBX_MAXLEV], range[GVDBX_MAXLEV][3], cnt0[GVDBX_MAXLEV], w0[GVDBX_MAXLEV];
    int cnt1[GVDBX_MAXLEV], w1[GVDBX_MAXLEV];
    for (int l = 0; l < levels; l++) {
        f.r(&ld[l], 4), f.r(&res[l], 4), f.r(range[l], 12), f.r(&cnt0[l], 4), f.r(&w0[l], 4), f.r(&cnt1[l], 4);
        f.r(&w1[l], 4);
    }
    if (!f.ok) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated level table", path);
    for (int l = 0; l < levels; l++) {
        if (ld[l] < 1 || ld[l] > 10 || res[l] != (1 << ld[l]) || cnt0[l] < 0 || cnt1[l] < 0)
            FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: bad level %d", path, l);
        if (w0[l] != (int)sizeof(GNode))   // VG.cpp:614-618
            FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: node records of %d bytes (this library: %d)", path, w0[l], (int)sizeof(GNode));
        if (l == 0 ? cnt1[0] != 0 : (w1[l] != 8 * res[l] * res[l] * res[l]))
            FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: child lists of level %d", path, l);
    }
    const int depth = (int)((root >> 8) & 0xFF);
    if ((root & 0xFF) != 0 || (root >> 16) != 0 || depth < 1 || depth >= levels || cnt0[depth] < 1 || cnt1[depth] < 1)
        FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: root id 0x%llx", path, (unsigned long long)root);
    TRY(gvdbx_configure(v, levels, ld));
    if (leafdim[0] != res[0] || leafdim[1] != res[0] || leafdim[2] != res[0] || leafcnt != cnt0[0])
        FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: brick count or size disagrees with the topology", path);
    // pools (PoolRead AL.cpp:985-993)
    std::vector<std::vector<GNode>> hn((size_t)levels);
    std::vector<std::vector<uint64_t>> hl((size_t)levels);
    for (int l = 0; l < levels; l++) {
        hn[l].resize((size_t)cnt0[l]);
        f.r(hn[l].data(), (size_t)cnt0[l] * sizeof(GNode));
    }
    for (int l = 1; l < levels; l++) {
        hl[l].resize((size_t)cnt1[l] * (size_t)(w1[l] / 8));
        f.r(hl[l].data(), (size_t)cnt1[l] * (size_t)w1[l]);
    }
    if (!f.ok) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated pools", path);
    // the topology state of a rebuild: counts, root, top level, and the
    // cell -> node maps of the nodes reached from the top node
    int64_t cnt[GVDBX_MAXLEV] = {};
    for (int l = 0; l < levels; l++) cnt[l] = cnt0[l];
    int tlev = 1;   // PrepareVDBPartially VG.cpp:3996-4012
    for (int l = levels - 1; l >= 0; l--)
        if (cnt[l] == 1) tlev = l;
    if (tlev < 1 || tlev > depth) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: top level %d below the leaves' parents", path, tlev);
    std::vector<std::vector<int32_t>> reach((size_t)levels);
    reach[tlev].push_back(0);
    for (int l = tlev; l >= 1; l--) {
        const int64_t len = (int64_t)res[l] * res[l] * res[l];
        for (int32_t k : reach[l]) {
            const uint64_t lid = hn[l][k].mChildList;
            if (lid == GVDBX_ID_UNDEFL || (int)((lid >> 8) & 0xFF) != l) continue;
            const int64_t li = (int64_t)(lid >> 16);
            if (li >= cnt1[l]) continue;
            for (int64_t b = 0; b < len; b++) {
                const uint64_t e = hl[l][(size_t)(li * len + b)];
                if (e == GVDBX_ID_UNDEF64) continue;
                const int64_t c = (int64_t)(int32_t)(uint32_t)(e >> 16);
                if ((int)((e >> 8) & 0xFF) != l - 1 || c < 0 || c >= cnt[l - 1])
                    FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: child list entry out of range at level %d", path, l);
                reach[l - 1].push_back((int32_t)c);
            }
        }
    }
    if (reach[0].empty()) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: no leaf is reachable", path);
    CellGrid grid[GVDBX_MAXLEV] = {};
    for (int l = 0; l < depth; l++) {
        if (reach[l].empty()) continue;
        int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
        for (int32_t k : reach[l])
            for (int a = 0; a < 3; a++) {
                const int c = (int)(hn[l][k].mPos[a] / v->range[l]);
                lo[a] = std::min(lo[a], c);
                hi[a] = std::max(hi[a], c);
            }
        grid[l] = CellGrid{lo[0], lo[1], lo[2], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    }
    // dense grids over the nodes' boxes, or sparse ones listing the node cells
    const bool sparse = v->topo_mode == GVDBX_TOPO_RADIX ||
                        (v->topo_mode == GVDBX_TOPO_AUTO && grid[0].count() > MAX_DENSE_CELLS);
    if (!sparse && grid[0].count() > MAX_DENSE_CELLS)
        FAIL(GVDBX_ERR_RANGE, "LoadVBX: %s: the leaves span %lld cells, beyond the dense-grid limit (topology mode DENSE)",
             path, (long long)grid[0].count());
    for (int l = 0; l < levels; l++) {
        TRY(ensure(v, v->nodes[l], (size_t)std::max<int64_t>(cnt[l], 1) * sizeof(GNode)));
        TRY(host_to_dev(v, v->nodes[l].p, hn[l].data(), (size_t)cnt[l] * sizeof(GNode)));
        if (l >= 1) {
            TRY(ensure(v, v->clist[l], std::max<size_t>(hl[l].size(), 1) * 8));
            TRY(host_to_dev(v, v->clist[l].p, hl[l].data(), hl[l].size() * 8));
        }
    }
    bool inactive = false;
    for (const GNode& n : hn[0]) inactive = inactive || !n.mFlags;
    CellGrid gb_sparse = {};
    for (int l = 0; l < depth && !sparse; l++) {
        std::vector<int32_t> idx((size_t)grid[l].count(), -1);
        for (int32_t k : reach[l]) {
            const GNode& n = hn[l][k];
            idx[(size_t)grid[l].dense_index((int)(n.mPos[0] / v->range[l]), (int)(n.mPos[1] / v->range[l]),
                                      (int)(n.mPos[2] / v->range[l]))] = k;
        }
        TRY(ensure(v, v->idx[l], std::max<size_t>(idx.size(), 1) * 4));
        TRY(host_to_dev(v, v->idx[l].p, idx.data(), idx.size() * 4));
    }
    if (sparse) {
        // per level the node cells in key order (a later node of one cell
        // replaces an earlier one, as the dense map does); the bin grid
        // lists the leaf cells dilated by one
        std::vector<uint64_t> leaf_keys;
        for (int l = 0; l < depth; l++) {
            CellGrid& g = grid[l];
            g.by = gvdbx_bits_for(g.ny);
            g.bz = gvdbx_bits_for(g.nz);
            std::vector<std::pair<uint64_t, int32_t>> kv;
            for (int32_t k : reach[l]) {
                const GNode& n = hn[l][k];
                kv.push_back({g.pack((int)(n.mPos[0] / v->range[l]), (int)(n.mPos[1] / v->range[l]),
                                     (int)(n.mPos[2] / v->range[l])), k});
            }
            std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            std::vector<uint64_t> keys;
            std::vector<int32_t> idx;
            for (size_t i = 0; i < kv.size(); i++) {
                if (i + 1 < kv.size() && kv[i + 1].first == kv[i].first) continue;
                keys.push_back(kv[i].first);
                idx.push_back(kv[i].second);
            }
            TRY(ensure(v, v->sg_keys[l], std::max<size_t>(keys.size(), 1) * 8));
            TRY(host_to_dev(v, v->sg_keys[l].p, keys.data(), keys.size() * 8));
            TRY(ensure(v, v->idx[l], std::max<size_t>(idx.size(), 1) * 4));
            TRY(host_to_dev(v, v->idx[l].p, idx.data(), idx.size() * 4));
            TRY(sg_table(v, v->sg_keys[l], (int64_t)keys.size(), v->sg_ht[l], g));
            if (l == 0) leaf_keys = keys;
        }
        const CellGrid& g0 = grid[0];
        CellGrid gb{g0.ox - 1, g0.oy - 1, g0.oz - 1, g0.nx + 2, g0.ny + 2, g0.nz + 2};
        gb.by = gvdbx_bits_for(gb.ny);
        gb.bz = gvdbx_bits_for(gb.nz);
        std::vector<uint64_t> bk;
        bk.reserve(leaf_keys.size() * 27);
        for (uint64_t k : leaf_keys) {
            int x, y, z;
            g0.unpack(k, x, y, z);
            for (int d = 0; d < 27; d++) bk.push_back(gb.pack(x + d / 9 - 1, y + (d / 3) % 3 - 1, z + d % 3 - 1));
        }
        std::sort(bk.begin(), bk.end());
        bk.erase(std::unique(bk.begin(), bk.end()), bk.end());
        TRY(ensure(v, v->sg_bkeys, std::max<size_t>(bk.size(), 1) * 8));
        TRY(host_to_dev(v, v->sg_bkeys.p, bk.data(), bk.size() * 8));
        TRY(sg_table(v, v->sg_bkeys, (int64_t)bk.size(), v->sg_bht, gb));
        gb_sparse = gb;
    }
    v->gb_sparse = gb_sparse;
    v->depth = depth;
    v->top_lev = tlev;
    for (int a = 0; a < 3; a++) v->root_pos[a] = hn[depth][0].mPos[a];
    for (int l = 0; l < GVDBX_MAXLEV; l++) {
        v->cnt[l] = l < levels ? cnt[l] : 0;
        v->nbase[l] = v->cnt[l];
        v->grid[l] = l < depth ? grid[l] : CellGrid{};
    }
    v->has_topo = true;
    v->has_atlas = false;
    v->topo_gen++;
    v->vis_valid = false;
    v->topo_mode_used = sparse ? GVDBX_TOPO_RADIX : GVDBX_TOPO_DENSE;
    v->leaf_inactive = inactive;
    v->force_full = false;
    v->ncloud = 0;
    TRY(compute_bounds(v));   // FinishTopology
    TRY(resolve_bounds(v));
    memcpy(v->pos_min, v->obj_min, sizeof(v->pos_min));   // no points were seen
    memcpy(v->pos_max, v->obj_max, sizeof(v->pos_max));
    // channels (DestroyChannels, AddChannel with the file's atlas size)
    for (int c = 0; c < GVDBX_MAXCHAN; c++) {
        free_buf(v->chan[c]);
        v->has_chan[c] = false;
    }
    v->clr_chan = -1;   // DestroyChannels VG.cpp:2474-2484
    v->atlas_cap = 0;
    std::vector<int> ctype((size_t)nch);
    for (int c = 0; c < nch; c++) {
        int type = 0, stride = 0;
        f.r(&type, 4), f.r(&stride, 4);
        if (!f.ok) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated channel %d", path, c);
        if (stride != 4) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: channel %d has %d-byte voxels", path, c, stride);
        TRY(gvdbx_add_channel(v, c, type, apron, 0, 0));
        ctype[c] = type;
        // the atlases follow one another; they are read after UpdateAtlas below
        const long skip = (long)axisres[0] * axisres[1] * axisres[2] * 4;
        if (fseek(f.fp, skip, SEEK_CUR) != 0) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated atlas", path);
    }
    if (nch == 0) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: no channel", path);
    const int br = brick_res(v);
    for (int a = 0; a < 3; a++)
        if (axiscnt[a] < 1 || axisres[a] != axiscnt[a] * br)
            FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: atlas of %dx%dx%d bricks, %d voxels wide", path, axiscnt[0], axiscnt[1],
                 axiscnt[2], axisres[0]);
    for (int a = 0; a < 3; a++) v->subdim[a] = axiscnt[a];
    TRY(update_atlas(v));
    // slot -> leaf of the active leaves, in leaf order (UpdateAtlas)
    const int64_t slots = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
    std::vector<int32_t> slot_leaf((size_t)slots, -1);
    int64_t rank = 0;
    for (int64_t k = 0; k < cnt[0] && rank < slots; k++)
        if (hn[0][k].mFlags) slot_leaf[(size_t)rank++] = (int32_t)k;
    // rewind to the first atlas and read them in order, file atlas slots as
    // UpdateAtlas assigns them
    const long per = (long)axisres[0] * axisres[1] * axisres[2] * 4 + 8;
    if (fseek(f.fp, -(long)nch * per, SEEK_CUR) != 0) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated atlas", path);
    const int64_t bv = brick_vox(v);
    const int64_t slice = (int64_t)axisres[0] * axisres[1];
    const int64_t fz = axisres[2] / br;   // brick layers in the file
    std::vector<uint32_t> bricks((size_t)(cnt[0] * bv), 0u), buf((size_t)slice);
    for (int c = 0; c < nch; c++) {
        int hdr[2];
        f.r(hdr, 8);
        std::fill(bricks.begin(), bricks.end(), 0u);
        for (int z = 0; z < axisres[2]; z++) {
            if (!f.r(buf.data(), (size_t)slice * 4)) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated atlas", path);
            const int bz = z / br, zi = z % br;
            if (bz >= fz) continue;
            for (int by = 0; by < axiscnt[1]; by++)
                for (int bx = 0; bx < axiscnt[0]; bx++) {
                    const int64_t s = ((int64_t)bz * axiscnt[1] + by) * axiscnt[0] + bx;
                    if (s >= slots || slot_leaf[(size_t)s] < 0) continue;
                    uint32_t* dst = bricks.data() + (size_t)slot_leaf[(size_t)s] * bv + (size_t)zi * br * br;
                    for (int y = 0; y < br; y++)
                        memcpy(dst + (size_t)y * br, buf.data() + (size_t)(by * br + y) * axisres[0] + (size_t)bx * br,
                               (size_t)br * 4);
                }
        }
        chan_written(v, c);
        TRY(host_to_dev(v, v->chan[c].p, bricks.data(), bricks.size() * 4));
    }
    (void)ctype;
    return GVDBX_OK;
}

}  // namespace

extern "C" {

int gvdbx_set_transform(gvdbx_volume* v, const float* pretrans, const float* scale, const float* angs,
                        const float* trans) {
    TRY(check_vol(v));
    if (!pretrans || !scale || !angs || !trans) FAIL(GVDBX_ERR_ARG, "SetTransform: null vector");
    memcpy(v->xform, pretrans, 12);
    memcpy(v->xform + 3, angs, 12);
    memcpy(v->xform + 6, scale, 12);
    memcpy(v->xform + 9, trans, 12);
    return GVDBX_OK;
}

int gvdbx_get_transform(gvdbx_volume* v, float* pretrans, float* scale, float* angs, float* trans) {
    TRY(check_vol(v));
    if (!pretrans || !scale || !angs || !trans) FAIL(GVDBX_ERR_ARG, "GetTransform: null vector");
    memcpy(pretrans, v->xform, 12);
    memcpy(angs, v->xform + 3, 12);
    memcpy(scale, v->xform + 6, 12);
    memcpy(trans, v->xform + 9, 12);
    return GVDBX_OK;
}

int gvdbx_save_vbx(gvdbx_volume* v, const char* path) {
    TRY(check_vol(v));
    if (!path) FAIL(GVDBX_ERR_ARG, "SaveVBX: null path");
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "SaveVBX: not available for batched clouds");
    TRY(need_atlas(v, "SaveVBX"));
    int nch = 0;
    TRY(vbx_channels(v, "SaveVBX", &nch));
    if (nch == 0) FAIL(GVDBX_ERR_STATE, "SaveVBX: no channel");
    TRY(sync(v));
    const int levels = v->levs, br = brick_res(v);
    std::vector<std::vector<char>> hn((size_t)levels), hl((size_t)levels);
    for (int l = 0; l < levels; l++) {
        hn[l].resize((size_t)v->cnt[l] * sizeof(GNode));
        TRY(dev_to_host(v, hn[l].data(), v->nodes[l].p, hn[l].size()));
        if (l >= 1) {
            hl[l].resize((size_t)(v->cnt[l] * (int64_t)v->res[l] * v->res[l] * v->res[l]) * 8);
            TRY(dev_to_host(v, hl[l].data(), v->clist[l].p, hl[l].size()));
        }
    }
    std::vector<GAtlasNode> amap((size_t)v->amap_cnt);
    TRY(dev_to_host(v, amap.data(), v->amap.p, amap.size() * sizeof(GAtlasNode)));
    VbxIO f;
    f.fp = fopen(path, "wb");
    if (!f.fp) FAIL(GVDBX_ERR_ARG, "SaveVBX: cannot open %s for writing", path);
    const unsigned char major = 1, minor = 11;   // VG.cpp:31-32
    const int num_grids = 1;
    f.w(&major, 1), f.w(&minor, 1), f.w(v->xform, 48), f.w(&num_grids, 4);
    const uint64_t grid_off = 2 + 48 + 4 + 8;
    f.w(&grid_off, 8);
    char name[256];
    memset(name, 0, sizeof(name));
    const unsigned char dtype = 'f', comps = 1, compress = 0, topotype = 2, layout = 0;
    const float vsize[3] = {1, 1, 1};
    const int leafcnt = (int)v->cnt[0], r0 = v->res[0], leafdim[3] = {r0, r0, r0}, apron = v->apron, reuse = 0;
    const int axisres[3] = {v->subdim[0] * br, v->subdim[1] * br, v->subdim[2] * br};
    const uint64_t atlas_sz = (uint64_t)4 * axisres[0] * (uint64_t)axisres[1] * axisres[2];
    f.w(name, 256), f.w(&dtype, 1), f.w(&comps, 1), f.w(&compress, 1), f.w(vsize, 12), f.w(&leafcnt, 4);
    f.w(leafdim, 12), f.w(&apron, 4), f.w(&nch, 4), f.w(&atlas_sz, 8), f.w(&topotype, 1), f.w(&reuse, 4);
    f.w(&layout, 1), f.w(v->subdim, 12), f.w(axisres, 12);
    const uint64_t root = (uint64_t)v->depth << 8;   // Elem(0, depth, 0)
    f.w(&levels, 4), f.w(&root, 8);
    for (int l = 0; l < levels; l++) {
        const int rg = (int)v->range[l], range[3] = {rg, rg, rg}, c0 = (int)v->cnt[l], w0 = (int)sizeof(GNode);
        const int c1 = l == 0 ? 0 : (int)v->cnt[l];   // one child list per node above the leaves
        const int w1 = l == 0 ? 0 : 8 * v->res[l] * v->res[l] * v->res[l];
        f.w(&v->logdim[l], 4), f.w(&v->res[l], 4), f.w(range, 12), f.w(&c0, 4), f.w(&w0, 4), f.w(&c1, 4), f.w(&w1, 4);
    }
    for (int l = 0; l < levels; l++) f.w(hn[l].data(), hn[l].size());
    for (int l = 1; l < levels; l++) f.w(hl[l].data(), hl[l].size());
    const int64_t bv = brick_vox(v), slice = (int64_t)axisres[0] * axisres[1];
    std::vector<uint32_t> bricks((size_t)(v->cnt[0] * bv)), buf((size_t)slice);
    for (int c = 0; c < nch; c++) {
        TRY(dev_to_host(v, bricks.data(), v->chan[c].p, bricks.size() * 4));
        const int type = v->chan_type[c], stride = 4;
        f.w(&type, 4), f.w(&stride, 4);
        for (int z = 0; z < axisres[2]; z++) {
            const int bz = z / br, zi = z % br;
            std::fill(buf.begin(), buf.end(), 0u);
            for (int by = 0; by < v->subdim[1]; by++)
                for (int bx = 0; bx < v->subdim[0]; bx++) {
                    const int64_t s = ((int64_t)bz * v->subdim[1] + by) * v->subdim[0] + bx;
                    if (s >= (int64_t)amap.size()) continue;
                    const int32_t leaf = amap[(size_t)s].mLeafID;
                    if (leaf < 0 || leaf >= v->cnt[0]) continue;
                    const uint32_t* src = bricks.data() + (size_t)leaf * bv + (size_t)zi * br * br;
                    for (int y = 0; y < br; y++)
                        memcpy(buf.data() + (size_t)(by * br + y) * axisres[0] + (size_t)bx * br, src + (size_t)y * br,
                               (size_t)br * 4);
                }
            f.w(buf.data(), (size_t)slice * 4);
        }
    }
    if (!f.ok) FAIL(GVDBX_ERR_HIP, "SaveVBX: write to %s failed", path);
    if (fclose(f.fp) != 0) {
        f.fp = nullptr;
        FAIL(GVDBX_ERR_HIP, "SaveVBX: closing %s failed", path);
    }
    f.fp = nullptr;
    return GVDBX_OK;
}

int gvdbx_load_vbx(gvdbx_volume* v, const char* path) { return gvdbx_load_vbx_version(v, path, 0, 0); }

int gvdbx_load_vbx_version(gvdbx_volume* v, const char* path, int force_maj, int force_min) {
    TRY(check_vol(v));
    if (!path) FAIL(GVDBX_ERR_ARG, "LoadVBX: null path");
    if (force_maj < 0 || force_maj > 255 || force_min < 0 || force_min > 255)
        FAIL(GVDBX_ERR_ARG, "LoadVBX: forced version %d.%d out of range", force_maj, force_min);
    VbxIO f;
    f.fp = fopen(path, "rb");
    if (!f.fp) FAIL(GVDBX_ERR_ARG, "LoadVBX: unable to open %s", path);
    unsigned char major = 0, minor = 0;
    f.r(&major, 1), f.r(&minor, 1);
    if (force_maj > 0 || force_min > 0) {   // VG.cpp:551-554: read the file as that version
        major = (unsigned char)force_maj;
        minor = (unsigned char)force_min;
    }
    float xf[12] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0};
    if ((major == 1 && minor >= 11) || major > 1) f.r(xf, 48);   // VG.cpp:562-569
    int num_grids = 0;
    f.r(&num_grids, 4);
    unsigned char masks = (major == 1 && minor == 0) ? 1 : 0;   // VG.cpp:577-588
    if (major >= 2) f.r(&masks, 1);
    if (!f.ok || num_grids < 1 || num_grids > 4096) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: bad header", path);
    if (masks) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: bitmask node records are not supported", path);
    std::vector<uint64_t> offs((size_t)num_grids);
    f.r(offs.data(), offs.size() * 8);
    if (!f.ok) FAIL(GVDBX_ERR_ARG, "LoadVBX: %s: truncated grid table", path);
    memcpy(v->xform, xf, sizeof(xf));
    for (int g = 0; g < num_grids; g++) TRY(vbx_load_grid(v, f, path));   // each grid replaces the last
    return GVDBX_OK;
}

}  // extern "C"

// ---------------- frame graphs (hipGraph of one level-set frame) ----------------
struct gvdbx_graph {
    gvdbx_volume* v = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t topo_gen = 0, buf_gen = 0, pos_gen = 0;
    const char* pos_base = nullptr;
    DevBuf sorted, sorted2;   // the binning buffers as the captured sort left them
};

extern "C" {

int gvdbx_frame_graph_create(gvdbx_volume* v, int num_pnts, int sc, float sc_radius, float radius, int chan,
                             int accum, int apron_chan, float boundval, gvdbx_graph** out) {
    TRY(check_vol(v));
    if (!out) FAIL(GVDBX_ERR_ARG, "FrameGraph: null output");
    *out = nullptr;
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "FrameGraph: not available for batched clouds");
    if (v->clr) FAIL(GVDBX_ERR_ARG, "FrameGraph: point colours are not captured (SetPoints without clr)");
    if (apron_chan >= 0) TRY(check_chan(v, apron_chan));
    // one eager frame: validates the arguments and sizes every buffer
    int64_t len = 0;
    TRY(insert_subcell(v, sc, num_pnts, sc_radius, 0.0f, 0.0f, 0.0f, &len, false));
    TRY(gather(v, 1, sc, radius, len, chan, -1, accum));
    if (apron_chan >= 0) TRY(update_apron(v, apron_chan, boundval));
    TRY(sync(v));
    // the same frame captured on a stream of its own
    hipStream_t cs = nullptr;
    GVDBX_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipStream_t keep = v->stream;
    const uint64_t bg = v->buf_gen;
    v->stream = cs;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int st = e == hipSuccess ? GVDBX_OK : GVDBX_ERR_HIP;
    if (st == GVDBX_OK) {
        const float t0[3] = {0.0f, 0.0f, 0.0f};
        v->bins_force = true;
        st = ensure_bins(v, num_pnts, t0, false);
        v->bins_force = false;
        if (st == GVDBX_OK) st = gather(v, 1, sc, radius, len, chan, -1, accum);
        if (st == GVDBX_OK && apron_chan >= 0) st = update_apron(v, apron_chan, boundval);
    }
    hipGraph_t graph = nullptr;
    if (e == hipSuccess) {
        const hipError_t e2 = hipStreamEndCapture(cs, &graph);
        if (st == GVDBX_OK && e2 != hipSuccess) {
            gvdbx_set_error("FrameGraph: capture failed: %s", hipGetErrorString(e2));
            st = GVDBX_ERR_HIP;
        }
    } else {
        gvdbx_set_error("FrameGraph: hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    v->stream = keep;
    (void)hipStreamDestroy(cs);
    if (st == GVDBX_OK && v->buf_gen != bg) {
        gvdbx_set_error("FrameGraph: a buffer was reallocated during capture");
        st = GVDBX_ERR_STATE;
    }
    hipGraphExec_t exec = nullptr;
    if (st == GVDBX_OK) {
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            gvdbx_set_error("FrameGraph: hipGraphInstantiate: %s", hipGetErrorString(e));
            st = GVDBX_ERR_HIP;
        }
    }
    if (st != GVDBX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return st;
    }
    gvdbx_graph* g = new gvdbx_graph;
    g->v = v;
    g->graph = graph;
    g->exec = exec;
    g->topo_gen = v->topo_gen;
    g->buf_gen = v->buf_gen;
    g->pos_gen = v->pos_gen;
    g->pos_base = v->pos.base;
    g->sorted = v->sorted;
    g->sorted2 = v->sorted2;
    v->graphs.push_back(g);
    *out = g;
    return GVDBX_OK;
}

int gvdbx_frame_graph_launch(gvdbx_graph* g) {
    if (!g) FAIL(GVDBX_ERR_ARG, "FrameGraph: null graph");
    if (!g->v) FAIL(GVDBX_ERR_STATE, "FrameGraph: its volume was destroyed");
    gvdbx_volume* v = g->v;
    TRY(check_vol(v));
    if (v->topo_gen != g->topo_gen || v->buf_gen != g->buf_gen || v->pos.base != g->pos_base)
        FAIL(GVDBX_ERR_STATE, "FrameGraph: the topology, a buffer or the point buffer changed since capture");
    GVDBX_CHECK(hipGraphLaunch(g->exec, v->stream));
    // the bins now hold the replayed points, in the buffers the capture used
    v->sorted = g->sorted;
    v->sorted2 = g->sorted2;
    v->bins_valid = true;
    v->sc_valid = true;
    return GVDBX_OK;
}

int gvdbx_frame_graph_destroy(gvdbx_graph* g) {
    if (!g) return GVDBX_OK;
    if (gvdbx_volume* v = g->v) {
        (void)hipStreamSynchronize(v->stream);
        v->graphs.erase(std::remove(v->graphs.begin(), v->graphs.end(), g), v->graphs.end());
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return GVDBX_OK;
}

}  // extern "C"

namespace {

// gvdbx_destroy: the volume's graphs lose their executables and their
// volume; their handles stay valid for gvdbx_frame_graph_destroy
void retire_graphs(gvdbx_volume* v) {
    for (gvdbx_graph* g : v->graphs) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        g->exec = nullptr;
        g->graph = nullptr;
        g->v = nullptr;
    }
    v->graphs.clear();
}

}  // namespace

// ---------------- autotuner (tile shapes of the result-invariant knobs) ----------------
namespace {

struct GatherCfg {
    int variant, classes, split, xcd;
};

int checksum_dev(gvdbx_volume* v, const void* p, int64_t words, unsigned long long out[2]) {
    TRY(ensure(v, v->misc, 256));
    unsigned long long* d = (unsigned long long*)((char*)v->misc.p + 128);
    launch_checksum((const uint32_t*)p, words, d, v->stream);
    TRY(check_launch("checksum"));
    return dev_to_host(v, out, d, 16);
}

// elapsed ms of `reps` calls of fn on the volume's stream (HIP events)
template <class F>
int time_calls(gvdbx_volume* v, int reps, F fn, float* ms) {
    hipEvent_t e0, e1;
    GVDBX_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        FAIL(GVDBX_ERR_HIP, "hipEventCreate failed");
    }
    int st = GVDBX_OK;
    (void)hipEventRecord(e0, v->stream);
    for (int r = 0; r < reps && st == GVDBX_OK; r++) st = fn();
    (void)hipEventRecord(e1, v->stream);
    if (st == GVDBX_OK && hipEventSynchronize(e1) == hipSuccess) {
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, e0, e1);
        *ms = t / (float)reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

}  // namespace

extern "C" {

int gvdbx_autotune_gather(gvdbx_volume* v, int mode, int sc, float radius, int chan, int reps,
                          gvdbx_tune_gather* out) {
    TRY(check_vol(v));
    if (!out || reps < 1 || (mode != 0 && mode != 1)) FAIL(GVDBX_ERR_ARG, "AutotuneGather: bad arguments");
    if (v->clr) FAIL(GVDBX_ERR_ARG, "AutotuneGather: tune without point colours");
    const char* what = "AutotuneGather";
    TRY(need_atlas(v, what));
    if (!v->sc_valid) FAIL(GVDBX_ERR_STATE, "%s: InsertPointsSubcell has not been called", what);
    const GatherCfg keep{v->gather_variant, v->gather_classes, v->gather_split, v->xcd_remap};
    auto apply = [&](const GatherCfg& c) {
        v->gather_variant = c.variant;
        v->gather_classes = c.classes;
        v->gather_split = c.split;
        v->xcd_remap = c.xcd;
    };
    auto run = [&]() -> int { return gather(v, mode, sc, radius, v->sc_len, chan, -1, 0); };
    const int64_t words = v->cnt[0] * brick_vox(v);
    // the configuration in place is the reference every candidate must reproduce
    TRY(run());
    unsigned long long ref[2];
    TRY(checksum_dev(v, v->chan[chan].p, words, ref));
    float ms0 = 0.0f;
    TRY(time_calls(v, reps, run, &ms0));
    std::vector<GatherCfg> cand;
    for (int x = 0; x <= 1; x++) {
        for (int sp : {0, 8, 16, 32, 64}) cand.push_back(GatherCfg{1, 1, sp, x});
        for (int sp : {0, 32, 64, 128}) cand.push_back(GatherCfg{1, 0, sp, x});
        cand.push_back(GatherCfg{0, 1, 0, x});
        cand.push_back(GatherCfg{2, 1, 0, x});
    }
    GatherCfg best = keep;
    float best_ms = ms0;
    int tried = 0;
    for (const GatherCfg& c : cand) {
        apply(c);
        int st = run();   // warm-up: sizes this candidate's buffers
        unsigned long long sum[2] = {0, 0};
        if (st == GVDBX_OK) st = checksum_dev(v, v->chan[chan].p, words, sum);
        if (st == GVDBX_OK && (sum[0] != ref[0] || sum[1] != ref[1])) {
            apply(keep);
            FAIL(GVDBX_ERR_STATE, "%s: variant %d classes %d split %d xcd %d changes the result", what, c.variant,
                 c.classes, c.split, c.xcd);
        }
        float ms = 0.0f;
        if (st == GVDBX_OK) st = time_calls(v, reps, run, &ms);
        if (st != GVDBX_OK) {
            apply(keep);
            return st;
        }
        tried++;
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    apply(best);
    TRY(run());   // leave the channel as the chosen configuration writes it (the same bits)
    out->gather_variant = best.variant;
    out->gather_classes = best.classes;
    out->gather_split = best.split;
    out->xcd_remap = best.xcd;
    out->best_ms = best_ms;
    out->default_ms = ms0;
    out->candidates = tried;
    return GVDBX_OK;
}

int gvdbx_autotune_binning(gvdbx_volume* v, int num_pnts, int sc, float radius, int reps, gvdbx_tune_binning* out) {
    TRY(check_vol(v));
    if (!out || reps < 1) FAIL(GVDBX_ERR_ARG, "AutotuneBinning: bad arguments");
    const char* what = "AutotuneBinning";
    if (v->ncloud > 0) FAIL(GVDBX_ERR_STATE, "%s: not available for batched clouds", what);
    TRY(need_atlas(v, what));
    const int keep_rec = v->bin_records, keep_tiles = v->sort_tiles;
    int64_t len0 = 0;
    auto run = [&]() -> int {
        v->bins_valid = false;
        int64_t len = 0;
        TRY(insert_subcell(v, sc, num_pnts, radius, 0.0f, 0.0f, 0.0f, &len, false));
        return len == len0 ? GVDBX_OK : GVDBX_ERR_STATE;
    };
    auto bins_sum = [&](unsigned long long s[2]) -> int {
        unsigned long long a[2], b[2];
        int32_t total = 0;   // binned points (those outside the bin grid sort last, unbinned)
        TRY(dev_to_host(v, &total, v->bin_off.as<int32_t>() + v->gbin.count(), 4));
        TRY(checksum_dev(v, v->bin_off.p, v->gbin.count() + 1, a));
        TRY(checksum_dev(v, v->sorted.p, (int64_t)total * 4, b));
        s[0] = a[0] ^ (b[0] * 31);
        s[1] = a[1] ^ (b[1] * 31);
        return GVDBX_OK;
    };
    v->bins_valid = false;
    TRY(insert_subcell(v, sc, num_pnts, radius, 0.0f, 0.0f, 0.0f, &len0, false));
    unsigned long long ref[2];
    TRY(bins_sum(ref));
    float ms0 = 0.0f;
    TRY(time_calls(v, reps, run, &ms0));
    int best_rec = keep_rec, best_tiles = keep_tiles, tried = 0;
    float best_ms = ms0;
    for (int rec : {1, 0})
        for (int tiles : {1, 2}) {
            v->bin_records = rec;
            v->sort_tiles = tiles;
            int st = run();
            unsigned long long sum[2] = {0, 0};
            if (st == GVDBX_OK) st = bins_sum(sum);
            if (st == GVDBX_OK && (sum[0] != ref[0] || sum[1] != ref[1])) st = GVDBX_ERR_STATE;
            float ms = 0.0f;
            if (st == GVDBX_OK) st = time_calls(v, reps, run, &ms);
            if (st != GVDBX_OK) {
                v->bin_records = keep_rec;
                v->sort_tiles = keep_tiles;
                v->bins_valid = false;
                FAIL(st, "%s: records %d tiles %d changes the bins or the list length", what, rec, tiles);
            }
            tried++;
            if (ms < best_ms) {
                best_ms = ms;
                best_rec = rec;
                best_tiles = tiles;
            }
        }
    v->bin_records = best_rec;
    v->sort_tiles = best_tiles;
    TRY(run());
    out->bin_records = best_rec;
    out->sort_tiles = best_tiles;
    out->best_ms = best_ms;
    out->default_ms = ms0;
    out->candidates = tried;
    return GVDBX_OK;
}

}  // extern "C"

