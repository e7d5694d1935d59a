// gvdbx_kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include "gvdbx_internal.h"

struct U2 {
    uint32_t a, b;
};

// ---------------- scan (exclusive prefix over cells) ----------------
// Number of 4096-element tiles of a scan over n items.
int64_t scan_tiles(int64_t n);

// ---------------- topology ----------------
void launch_bbox_init(BBoxAccum* acc, hipStream_t s);
// parts: bbox_parts(n) scratch partials
int bbox_parts(int64_t n);
void launch_bbox(PointSrc src, int64_t n, float tx, float ty, float tz, BBoxAccum* acc, BBoxAccum* parts,
                 hipStream_t s);

struct ReachParams {
    CellGrid g0;
    int range0;
    int logr0;
    int use_mask;      // 1: 3x3x3 reach mask per cell; 0: no extra bricks (radius < 1)
    float radius;
    CloudMap cm;       // batched clouds (ncloud 0: none)
};
// generic path for radii beyond half a brick: mark every box cell directly
void launch_mark_boxes(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid g0, int range0,
                       float radius, uint8_t* all0, hipStream_t s);
// base0 = reach0 != 0; all0 |= base0 and every reached neighbour
// allowed (optional): per g0 cell, >= 0 where extra bricks may be created
void launch_expand_reach(CellGrid g0, const uint32_t* reach0, uint8_t* base0, uint8_t* all0, int generic,
                         const int32_t* allowed, hipStream_t s);
// level l flags from level l-1 flags (every flagged child marks its parent;
// level l's flags must be clear)
void launch_level_up(CellGrid gc, const uint8_t* base_c, const uint8_t* all_c, CellGrid gp, int logres_p,
                     uint8_t* base_p, uint8_t* all_p, hipStream_t s);

// ordered scan of (base, extra) flags of one level; partials must hold
// scan_tiles(n)+1 entries, totals gets (#base, #extra)
void launch_flag_scan_reduce(const uint8_t* base, const uint8_t* all, int64_t n, U2* partials, U2* totals,
                             hipStream_t s);
struct NodeBuildParams {
    CellGrid g;
    int lev;
    int64_t range;
    int is_leaf;
    GNode* nodes;
    int32_t* idx;        // cell -> node index or -1
    int64_t node_off;    // pool index of the first node made (appending to a pool)
    int keep_idx;        // 1: idx is written only where a node is made
};
// flag scan apply of up to GVDBX_MAXLEV levels in one launch (tile_start is filled in)
struct FlagApplyLevel {
    const uint8_t* base;
    const uint8_t* all;
    int64_t n;
    const U2* partials;
    const U2* totals;
    NodeBuildParams np;
};
struct FlagApplyBatch {
    FlagApplyLevel lv[GVDBX_MAXLEV];
    int64_t tile_start[GVDBX_MAXLEV + 1];
    int n;
};
void launch_flag_scan_apply(FlagApplyBatch& fb, hipStream_t s);
// a[0..na) = 0 and b[0..nb) = 0 in one launch
// clears up to 2 * GVDBX_MAXLEV + 2 buffers (sizes in bytes) in one launch
struct ZeroBatch {
    uint8_t* p[2 * GVDBX_MAXLEV + 2];
    int64_t bytes[2 * GVDBX_MAXLEV + 2];
    int n;
};
void launch_zero_batch(const ZeroBatch& zb, hipStream_t s);
void launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s);
// out[0] = sum of p[i], out[1] = sum of p[i] * (i mod 65521 + 1) (mod 2^64)
void launch_checksum(const uint32_t* p, int64_t n, unsigned long long* out, hipStream_t s);
void launch_fill_f32(float* p, int64_t n, float v, hipStream_t s);
void launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);

struct LinkParams {
    int lev;              // child level
    int64_t nchild;
    const GNode* child;
    GNode* childw;
    const GNode* parent_nodes;
    uint64_t* parent_clist;
    int res_p;            // res of level lev+1
    int64_t range_p;      // range of level lev+1
    int logres_p;
    int parent_is_root;
    CellGrid gp;          // grid of level lev+1 (unused for the root)
    const int32_t* idx_p;
    int64_t nparent;
    int need_flag;        // 1: only nodes with mFlags set are linked (gvdbLinkBricks)
};
// the links of up to GVDBX_MAXLEV levels: level i links children [start[i], start[i+1])
struct LinkBatch {
    LinkParams lv[GVDBX_MAXLEV];
    int64_t start[GVDBX_MAXLEV + 1];
    int n;
};
void launch_link(const LinkBatch& lb, hipStream_t s);
// child lists list[i][0 .. start[i+1]-start[i]) set to "no child"; root copied to root_dst (if set)
struct ClistInit {
    uint64_t* list[GVDBX_MAXLEV];
    int64_t start[GVDBX_MAXLEV + 1];
    int n;
    GNode root;
    GNode* root_dst;
};
void launch_clist_init(const ClistInit& ci, hipStream_t s);
void launch_bounds_init(int32_t* out6, hipStream_t s);

// leaf k -> mValue (closed form of getAtlasPos) and atlas map slot k
void launch_atlas_assign(GNode* leaves, int64_t nleaf, int sx, int sy, int brick_res, int apron,
                         GAtlasNode* amap, int64_t nmap, hipStream_t s);
void launch_leaf_bounds(const GNode* leaves, int64_t nleaf, int range0, int32_t* out6, hipStream_t s);

// ---------------- sparse cell grids (sparse.hip) ----------------
// packed home cell (in box) of every valid point, `none` otherwise
void launch_sp_point_keys(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid box, int logr0,
                          uint64_t none, uint64_t* keys, hipStream_t s);
// flag[i] = 1 at the head of each run of equal sorted keys other than `none`
void launch_sp_heads(const uint64_t* keys, int64_t n, uint64_t none, int32_t* flag, hipStream_t s);
// out[off[i]] = keys[i] where flag[i]
void launch_sp_compact(const uint64_t* keys, const int32_t* flag, const int32_t* off, int64_t n, uint64_t* out,
                       hipStream_t s);
// (hi-lo+1)^3 outputs per input cell: the cells at offsets [lo, hi]^3, packed in go or `none`
void launch_sp_dilate(const uint64_t* in, int64_t m, CellGrid gi, int lo, int hi, CellGrid go, uint64_t none,
                      uint64_t* out, hipStream_t s);
// parent (>> logres) of every input cell, packed in gp or `none`
void launch_sp_parents(const uint64_t* in, int64_t m, CellGrid gc, int logres, CellGrid gp, uint64_t none,
                       uint64_t* out, hipStream_t s);
// hash table of m distinct keys (rank i for keys[i]); ht pre-filled with ~0
void launch_sp_hash(const uint64_t* keys, int64_t m, uint4* ht, uint32_t hmask, hipStream_t s);

// ---------------- point binning ----------------
struct BinParams {
    CellGrid gb;          // binning grid (leaf grid grown by one cell)
    int range0;
    int logr0;
    int sc_shift;
    // subcell list length (InsertPointsSubcell out-parameter)
    int sc;
    float radius;
    CellGrid g0;
    const uint32_t* nbmask;     // per g0 cell: 27-bit neighbour-leaf mask (launch_leaf_nbmask)
};
// reach masks of level-0 cells from cell-sorted points (total = off[gb.count()])
void launch_reach_sorted(const float4* sorted, const int32_t* total, int64_t nmax, ReachParams rp, uint32_t* reach0,
                         hipStream_t s);
// InsertPointsSubcell list length over cell-sorted points
void launch_subcell_len(const float4* sorted, const int32_t* total, int64_t nmax, BinParams bp,
                        unsigned long long* len, hipStream_t s);
void launch_count_scan_reduce(const int32_t* cnt, int64_t n, uint32_t* partials, uint32_t* total, hipStream_t s);
void launch_count_scan_apply(const int32_t* cnt, int64_t n, const uint32_t* partials, int32_t* off,
                             hipStream_t s);
// also sets *reset = 0 when reset is non-null
void launch_leaf_nbmask(const int32_t* idx0, CellGrid g0, uint32_t* mask, unsigned long long* reset, hipStream_t s);

// ---------------- points -> voxels ----------------
struct GatherParams {
    const int32_t* leaf_cloud;   // batched clouds: cloud of each leaf (else null)
    const int32_t* cell_off;     // 3 per cloud, level-0 cells
    const GNode* leaves;
    int64_t nleaf;
    CellGrid gb;
    const int32_t* off;       // gb.count()+1 offsets into sorted
    const float4* sorted;
    float* chan;              // brick-linear atlas channel
    int brick_res;
    int apron;
    int mode;                 // 0 density (max), 1 level set (min)
    int accum;
    int sc;
    int sc_shift;             // log2(sc): sc divides the power-of-two brick
    float box_radius;         // radius given to InsertPointsSubcell
    float radius;             // radius given to Gather*
    float cutoff;             // voxels farther than this cannot change a value
    int variant;              // 0: LDS-staged candidates, 1: ballot/readlane walk
    int xcd_remap;            // 1: give each XCD a contiguous run of tiles
    const uint32_t* box;      // variant 1, optional: per bin entry, launch_gather_box's packed ranges
    const int32_t* coff;      // variant 1, optional: class starts of launch_gather_classes (sorted is then
                              // its reordered copy, packed ranges in .w; box unused)
    const uint8_t* cboth;     // with coff: per bin, axes along which a point reaches past both faces
    int quarters;             // 8-voxel tiles with classes: z pruned per 2-voxel quarter instead of per half
    int push;                 // 8^3 bricks, apron 1: also write boundary voxels into the neighbours' aprons
    const int32_t* bin_leaf;  // push: per gb cell, its leaf or -1 (launch_bin_leaf)
};
// per gb cell: the leaf at that cell (idx0 over g0, leaves < nleaf) or -1
void launch_bin_leaf(CellGrid gb, CellGrid g0, const int32_t* idx0, int64_t nleaf, int32_t* out, hipStream_t s);
// Per bin entry (the first *total), the membership range clipped to the
// cutoff along each axis, as offsets from floor(p): bits 10a..10a+4 hold
// floor(p) + 1 - lo, bits 10a+5..10a+9 hold hi - floor(p); bit 31: empty on
// some axis.  Needs ceil(box_radius) + sc <= 29 (every offset fits 5 bits).
void launch_gather_box(const float4* sorted, const int32_t* total, int64_t nmax, float box_radius, int sc, int sc_shift,
                       float cutoff, uint32_t* box, hipStream_t s);
// Per bin (level-0 cell of side res0 = 2^k), its entries reordered stably by
// reach class (bit a: the clipped range along axis a reaches past the bin's
// upper face) into csorted as (x, y, z, launch_gather_box's packed ranges),
// the class starts in coff (8 per bin plus the total) and in both[bin] the
// axes along which some entry reaches past both faces.  Same range limits
// as launch_gather_box.  Scratch for bins of more than 256 entries, which
// are cut into units of 256: ctr (2 int32), units (unit_cap int2, at least
// entries / 128 + 2), bigs (big_cap int4, at least entries / 256 + 1),
// ucnt (9 int32 per unit).  qz (optional): FP16 lists; the entries' xyz
// become their FP16 round trips (membership still from the inserted point),
// and the offsets need ceil(box_radius) + sc <= 28.
void launch_gather_classes(const float4* sorted, const int32_t* off, CellGrid gb, int res0, float box_radius, int sc,
                           int sc_shift, float cutoff, float4* csorted, int32_t* coff, uint8_t* both, int32_t* ctr,
                           int2* units, int64_t unit_cap, int4* bigs, int64_t big_cap, int32_t* ucnt, hipStream_t s,
                           const Quant16* qz = nullptr);
constexpr int GATHER_SPLIT_DEFAULT = 16;   // chunks per work item of a dense tile (0: one wave per tile)
void launch_gather(GatherParams gp, hipStream_t s);
// Variant 1 with tiles of more than 2 split chunks of 64 candidates cut into
// work items of about `split` chunks: items (item_cap int4), counts (2 int32),
// slot_tile and acc (slot_cap tiles of 512 uint32) are scratch.
void launch_gather_split(GatherParams gp, int split, int4* items, int64_t item_cap, int32_t* counts,
                         int32_t* slot_tile, int64_t slot_cap, uint32_t* acc, hipStream_t s);

struct SplatParams {
    const int32_t* leaf_cloud;   // batched clouds: cloud of each leaf (else null)
    const int32_t* cell_off;     // 3 per cloud, level-0 cells
    const GNode* leaves;
    int64_t nleaf;
    CellGrid gb;
    const int32_t* off;
    const float4* sorted;     // (x, y, z, weight) in bin order (launch_splat_src)
    const float4* bins;       // the bins themselves (x, y, z, point index)
    CloudMap cm;              // batched clouds (ncloud 0 otherwise)
    const char* weights;      // may be null
    int64_t wstride;
    float* chan;
    int brick_res;
    int apron;
    int accum;
    float radius;
};
// Cubic splat.  Cell push: a bin cell's points in runs of `chunk` (runs[c]
// per cell, run_off = exclusive scan, run_cell maps run -> cell), each run's
// (res0 + 2h)^3 grid in `part` (h = ceil(r)), collected per leaf.  Tile pull:
// one kernel, for grids that do not fit LDS.
constexpr int SPLAT_CHUNK_DEFAULT = 1024;
int splat_halo(float radius);
int64_t splat_cell_lds(int res0, float radius);   // bytes of LDS per run
void launch_splat_runs(SplatParams sp, int chunk, int32_t* runs, hipStream_t s);
void launch_splat_run_map(SplatParams sp, const int32_t* run_off, int32_t* run_cell, hipStream_t s);
void launch_splat_cell(SplatParams sp, const int32_t* run_cell, const int32_t* run_off, int nruns, int chunk,
                       float* part, hipStream_t s);
void launch_splat_collect(SplatParams sp, const int32_t* run_off, const float* part, hipStream_t s);
void launch_splat_tile(SplatParams sp, hipStream_t s);
// (x, y, z, weight) of the first *count bin-sorted points (weight 1 without
// weights); count is the device-side number of binned points.
void launch_splat_src(const float4* sorted, int64_t n, const int32_t* count, const char* weights, int64_t wstride,
                      float4* out, hipStream_t s);

struct ApronParams {
    const GNode* leaves;
    int64_t nleaf;
    CellGrid g0;
    const int32_t* idx0;
    int range0;
    int logr0;
    float* chan;              // 4-byte voxels of either type (the apron copies words)
    int brick_res;
    float boundval;
    uint32_t bound_bits;      // the word stored outside every leaf (float bits or packed uchar4)
    int order;                // 8^3 bricks, dense g0: 1 = bricks taken in cell order with x fastest, 0 = leaf order
    int pushed;               // 8^3 bricks: the apron voxels that have a neighbour leaf are current (pushing
                              // gather); only the boundary word is written, where no leaf is
};
void launch_apron(ApronParams ap, hipStream_t s);

// ---------------- colour channels (uchar4 words) ----------------
struct ColorGatherParams {
    const GNode* leaves;
    int64_t nleaf;
    CellGrid gb;              // bins of the last InsertPointsSubcell
    const int32_t* off;
    const float4* sorted;     // (x, y, z, point index bits)
    const char* clr;          // uchar4 per point at clr + i * clr_stride
    int64_t clr_stride;
    uint32_t* chan;           // brick-linear colour channel
    int brick_res, apron, res0, logr0;
    int sc, sc_shift;
    float R;                  // InsertPointsSubcell radius
    int accum;
};
// colour mean of the gathers, one workgroup per leaf
void launch_gather_color(ColorGatherParams cp, hipStream_t s);
int color_max_subcells();
struct ColorScatterParams {
    const GNode* leaves;
    int64_t nleaf;
    CellGrid gb;
    const int32_t* off;
    const float4* sorted;
    const char* clr;
    int64_t clr_stride;
    uint32_t* chan;
    int brick_res, apron, res0, logr0;
    int avg;
};
// colour part of ScatterDensity, one workgroup per leaf
void launch_scatter_color(ColorScatterParams cp, hipStream_t s);

struct InsertParams {
    CellGrid gb;
    CellGrid g0;
    const int32_t* idx0;
    const int32_t* off;
    const float4* sorted;
    int logr0;
    int32_t* pnode;
    int32_t* poff;
    int64_t n;
};
void launch_insert_points(InsertParams ip, hipStream_t s);

struct ScatterParams {
    const GNode* leaves;
    int64_t nleaf;
    CellGrid gb;
    const int32_t* off;
    const float4* sorted;
    float* chan;
    int brick_res;
    int apron;
    int res0;
    float radius;
    int expand;
};
void launch_scatter_density(ScatterParams sp, hipStream_t s);

// ---------------- radix sort (LSD, 8-bit digits) ----------------
// Sorts n (key, value) pairs on key bits [begin_bit, end_bit), stable.
// Buffers: keys/vals hold the input; tkeys/tvals are scratch of n entries;
// hist needs radix_hist_entries(n) int32, off the same + 1, partials
// scan_tiles(radix_hist_entries(n) + 1) uint32.  Returns 1 when the result
// ended in tkeys/tvals (odd pass count), else 0.
int64_t radix_hist_entries(int64_t n);
// iota: vals holds no input; the values are the entries' indices 0..n-1
// (the first pass makes them), vals still serves as a buffer.  wide: tiles
// of twice the entries (longer runs per digit, fewer blocks per CU), for
// large sorts.
int radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* tkeys, uint32_t* tvals, int64_t n, int begin_bit,
                     int end_bit, int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s, bool iota = false);
int radix_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t* tkeys, uint32_t* tvals, int64_t n, int begin_bit,
                       int end_bit, int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s, bool iota = false,
                       bool wide = false);
void launch_iota_u32(uint32_t* p, int64_t n, hipStream_t s);

// ---------------- binning by radix sort ----------------
// Points binned by cell of gb (leaf grid grown by one cell) with a stable LSD
// radix sort that moves (x, y, z, index) records (translated positions;
// invalid points as (-1, -1, -1, index), sorting last with the points outside
// gb).  buf0 / buf1: n float4 each; returns which holds the result.  keys
// (n uint32) gets the sorted keys; with rp, each leaf cell's extra-brick
// reach goes to reach0.  hist / off / partials as for radix_sort_pairs32.
int bin_sort_records(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid gb, int logr0, float4* buf0,
                     float4* buf1, uint32_t* keys, int32_t* hist, int32_t* off, uint32_t* partials,
                     const ReachParams* rp, uint32_t* reach0, hipStream_t s);
// key[i] = bin cell of point i (gb.count() for invalid / outside points),
// val[i] = i
void launch_bin_keys(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid gb, int logr0, CloudMap cm,
                     uint32_t* key, uint32_t* val, hipStream_t s);   // val may be null
// sorted[j] = (w of point val[j], val[j]) for keys < ncell; with rp, also
// the extra-brick reach of those points into reach0 (launch_reach_sorted's work)
void launch_bin_gather(PointSrc src, int64_t n, float tx, float ty, float tz, const uint32_t* key,
                       const uint32_t* val, uint32_t ncell, float4* sorted, const ReachParams* rp, uint32_t* reach0,
                       hipStream_t s);
// off[c] = first sorted position with key >= c, c in [0, ncell]
void launch_bin_offsets(const uint32_t* key, int64_t n, int64_t ncell, int32_t* off, hipStream_t s);

// ---------------- rendering ----------------
struct RenderTree {
    const GNode* nodes[GVDBX_MAXLEV];
    const uint64_t* clist[GVDBX_MAXLEV];
    int64_t ncnt[GVDBX_MAXLEV];
    int logdim[GVDBX_MAXLEV];
    int res[GVDBX_MAXLEV];
    float vdel[GVDBX_MAXLEV];   // child range of level l (range of level l-1)
    int levs, top_lev;
    float bmin[3], bmax[3];
    const float* chan;          // brick-linear padded atlas channel
    int brick_res;
};
struct RenderCam {
    int width, height;
    float org[3], org_u[3], org_v[3];
    float dir[3], dir_u[3], dir_v[3];
    // batched clouds: the camera's cloud sits shifted by `shift` voxels in the
    // packed tree, and its rays are bounded by its own leaf bounds
    int shift[3];
    int has_bounds;
    float bmin[3], bmax[3];
};
struct RenderScene {
    float step, extinct, albedo, minval, alphacut, thresh, vmin, vmax, epsilon;
    int max_iter;
    float backclr[4];
    float light[3];
    float shadow_amt, shadow_bias;
    const float4* lut;
    float finestep;                      // SCN_FINESTEP (SHADE_TRICUBIC)
    float slice_pnt[3], slice_norm[3];   // SCN_SLICE_PNT / SCN_SLICE_NORM (sections)
    const uchar4* clr;                   // SetColorChannel: the uchar4 channel colouring Render, or null
};
// per (ray, brick) segment entry state of a recorded EA raycast
// Segment checkpoints of a recording EA raycast.  Each wave buffers its
// checkpoints and writes them out in chunks; a chunk is described by the tag
// (global wave index << 32 | the wave's chunk sequence number), its first
// record and its size, so that the records can later be put in an order
// that does not depend on which wave wrote first.
struct VisitRec {
    uint32_t* key;       // leaf of each record
    float4* rec;         // (t handed to the brick function, T and I at entry, ray id bits)
    uint32_t* count;     // count[0]: records, count[1]: chunks, one 8-byte-aligned pair (may exceed the caps: then rerun)
    int64_t cap;
    uint64_t* chunk_tag;   // (global wave index << 32) | records the wave wrote in earlier chunks
    uint32_t* chunk_base;
    uint32_t* chunk_size;
    int64_t chunk_cap;
    uint32_t* wave_recs;   // per global wave: records written (zero-filled by the host)
};
void launch_render(int shade, RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba, uchar4* rgba8,
                   hipStream_t s);
// ScnRay (cuda_gvdb_scene.cuh:66-74): 64-byte ray record of Raytrace
struct GScnRay {
    float hit[3], normal[3], orig[3], dir[3];
    uint32_t clr, pnode, pndx, pad;
};
static_assert(sizeof(GScnRay) == 64, "ray record must be 64 bytes");
void launch_raytrace(RenderTree tv, RenderScene sc, GScnRay* rays, int64_t n, float bias, hipStream_t s);
// cams (device, ncam entries, all of cam's size) renders one image per
// camera, image c at c*width*height; cams == nullptr renders cam alone
void launch_raycast_ea(RenderTree tv, RenderCam cam, const RenderCam* cams, int ncam, RenderScene sc, float* I,
                       float* T, int record, VisitRec vr, hipStream_t s);

struct EABackParams {
    RenderTree tv;
    RenderCam cam;
    const RenderCam* cams;  // batched: camera of ray r is cams[r / (width*height)]
    RenderScene sc;
    const uint32_t* idx;    // record index of each segment, grouped by leaf in a fixed order
    const float4* rec;      // (t, T, I, ray id bits)
    const int32_t* seg;     // per leaf [begin, end) into idx, 2 per leaf
    const float* I;
    const float* T;
    const float* gI;
    const float* gT;        // may be null
    float* grad;            // padded bricks
};
// per leaf [begin, end) of the leaf-sorted keys
void launch_segments(const uint32_t* key, int64_t n, int32_t* seg, int64_t nleaf, hipStream_t s);
// chunk c (tag[c] = wave << 32 | records before it in its wave) is copied to
// wave_off[wave] + that offset: key_out = its records' leaves, val_out =
// their record indices
void launch_chunk_order(const uint64_t* tag, int64_t nchunk, const uint32_t* chunk_base, const uint32_t* chunk_size,
                        const int32_t* wave_off, const uint32_t* key, uint32_t* key_out, uint32_t* val_out,
                        hipStream_t s);
void launch_ea_backward(EABackParams p, hipStream_t s);

struct SplatBackParams {
    CloudMap cm;
    CellGrid g0;
    const int32_t* idx0;
    const GNode* leaves;
    int64_t nleaf;
    int logr0;
    int brick_res;
    const float4* sorted;      // bin-sorted points (x, y, z, original index bits)
    const int32_t* total;      // number of sorted points
    int64_t nmax;
    const char* weights;
    int64_t wstride;
    float radius;
    const float* grad;
    float* dpos;
    float* dw;
};
void launch_splat_backward(SplatBackParams p, hipStream_t s);
// cell-push adjoint over the runs of launch_splat_runs (bins gb / off)
void launch_splat_cell_backward(SplatBackParams p, CellGrid gb, const int32_t* off, const int32_t* run_cell,
                                const int32_t* run_off, int nruns, int res0, int chunk, hipStream_t s);
void launch_apron_adjoint(ApronParams ap, float* grad, hipStream_t s);

// ---------------- batched inputs ----------------
void launch_backproject(const float* depth, int nframes, int width, int height, const float* intr4,
                        const float* poses, float voxel, const float* origin3, float dmin, float dmax, float* out,
                        hipStream_t s);

// per-cloud bounding boxes over (cloud, chunk) tiles: tiles[2t] = cloud,
// tiles[2t+1] = first point of the chunk (chunks of BBOX_CHUNK points)
constexpr int BBOX_CHUNK = 8192;
void launch_bbox_clouds(PointSrc src, const int32_t* tiles, int64_t ntiles, const int64_t* pt_off, BBoxAccum* acc,
                        hipStream_t s);
void launch_bbox_clouds_init(BBoxAccum* acc, int ncloud, hipStream_t s);
// cell_cloud[g0 cell] = c over each cloud's box of cells (boxes 6 ints per
// cloud, packed cells, inclusive lo / exclusive hi); tiles as for bboxes over
// CELL_CHUNK cells of each box
constexpr int CELL_CHUNK = 4096;
void launch_paint_clouds(CellGrid g0, const int32_t* boxes, const int32_t* tiles, int64_t ntiles, int32_t* cell_cloud,
                         hipStream_t s);
// leaf_cloud[k] = cell_cloud[cell of leaf k]
void launch_leaf_clouds(const GNode* leaves, int64_t nleaf, CellGrid g0, int logr0, const int32_t* cell_cloud,
                        int32_t* leaf_cloud, hipStream_t s);
// per-cloud leaf bounds (6 ints per cloud, packed voxels): min mPos, max mPos + range0
void launch_leaf_bounds_clouds(const GNode* leaves, int64_t nleaf, int range0, const int32_t* leaf_cloud,
                               int32_t* out6, hipStream_t s);

// ---------------- incremental topology (ActivateIncreBricksGPU) ----------------
// old -> new grid: dst[c] = src[c] where the old grid holds cell c, else -1
void launch_remap_idx(CellGrid from, const int32_t* src, CellGrid to, int32_t* dst, hipStream_t s);
struct IncreMark {
    int depth;
    int64_t range[GVDBX_MAXLEV];
    int root_cell[GVDBX_MAXLEV][3];   // the root's first cell of each level
    int root_span[GVDBX_MAXLEV];      // cells of each level across the root
    CellGrid g[GVDBX_MAXLEV];         // grids that hold every cell the update can touch
    const int32_t* idx[GVDBX_MAXLEV]; // cell -> reachable node before the update, or -1
    uint8_t* newf[GVDBX_MAXLEV];      // out: 1 where a touched cell has no node
    uint8_t* marker0;                 // out: 1 for every reached leaf
    int64_t nleaf;
};
void launch_incre_mark(PointSrc src, int64_t n, float tx, float ty, float tz, float radius, const IncreMark& im,
                       hipStream_t s);
// leaf k < n: mFlags = (accum ? mFlags : 0) | marker[k]
void launch_incre_leaf_flags(GNode* leaves, int64_t n, const uint8_t* marker, int accum, hipStream_t s);
struct DelinkParams {
    int lev;
    GNode* nodes;
    int64_t n;
    const uint64_t* own_clist;   // lists of level lev (lev >= 1)
    int64_t own_len;             // res(lev)^3
    int parent_is_root;
    CellGrid gp;                 // grid of level lev+1
    const int32_t* idx_p;        // its reachable map before the update
    GNode* parent_nodes;
    uint64_t* parent_clist;
    int res_p;
    int64_t range_p;
    int64_t nparent;
};
void launch_incre_delink(const DelinkParams& dp, hipStream_t s);
// idx[c] = -1 where its node was delinked (mParent undefined)
void launch_incre_final_idx(CellGrid g, int32_t* idx, const GNode* nodes, hipStream_t s);
// flags[k] = leaves[k].mFlags != 0
void launch_leaf_flags32(const GNode* leaves, int64_t n, int32_t* flags, hipStream_t s);
// UpdateAtlas over active leaves: active leaf k takes atlas slot rank[k]
// (mValue, atlas map entry); inactive leaves keep their mValue; map slots
// no leaf takes are undefined
void launch_atlas_assign_rank(GNode* leaves, int64_t nleaf, const int32_t* rank, int sx, int sy, int brick_res,
                              int apron, GAtlasNode* amap, int64_t nmap, hipStream_t s);
