/*
 * gvdbx_c.h -- C-ABI of the MI355X-native point -> sparse-volume library.
 *
 * This is the drop-in boundary for GVDB Voxels' points path.  Each entry
 * point replaces one method of nvdb::VolumeGVDB; the citation after each
 * declaration names the reference method (paths relative to the reference
 * repository; VG.h = source/gvdb_library/src/gvdb_volume_gvdb.h, VG.cpp =
 * source/gvdb_library/src/gvdb_volume_gvdb.cpp).
 *
 * Conventions
 *   - every entry point returns a status (GVDBX_OK == 0); the message of the
 *     last failure on the calling thread is gvdbx_last_error();
 *   - a failing call leaves the volume as it was before the call;
 *   - pointers named *_dev are device pointers (HBM); *_host are host memory;
 *   - the library never frees caller memory;
 *   - work is issued on the volume's stream (gvdbx_set_stream; default: the
 *     null stream); calls that return host data synchronise that stream.
 */
#ifndef GVDBX_C_H
#define GVDBX_C_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gvdbx_volume gvdbx_volume;

enum gvdbx_status {
    GVDBX_OK = 0,
    GVDBX_ERR_ARG = 1,       /* bad argument or handle */
    GVDBX_ERR_STATE = 2,     /* call out of order (e.g. no topology / no channel) */
    GVDBX_ERR_DEPTH = 3,     /* point extents need depth 0 or more levels than configured */
    GVDBX_ERR_NOPOINTS = 4,  /* no valid point */
    GVDBX_ERR_RANGE = 5,     /* extents exceed key or grid limits */
    GVDBX_ERR_HIP = 6,       /* HIP runtime failure */
    GVDBX_ERR_NOMEM = 7      /* device allocation failed */
};

/* channel types, numbered as the reference's T_UCHAR4 / T_FLOAT
 * (source/gvdb_library/src/gvdb_types.h:75-76); both store 4 bytes per
 * voxel in the same brick layout (uchar4: x in the lowest byte) */
enum gvdbx_dtype { GVDBX_UCHAR4 = 2, GVDBX_FLOAT = 3 };

/* How RebuildTopology finds the cells of each tree level.  DENSE: per-cell
 * arrays over the points' box (fast; GVDBX_ERR_RANGE beyond 2^28 level-0
 * cells).  RADIX: the occupied cells only, found by an LSD radix sort of
 * 64-bit cell keys with run-head dedup (the reference's RadixSortByByte +
 * FindUnique path), looked up through a hash table; any box within the
 * 16-bit-per-axis key range.  AUTO: DENSE when the box fits, else RADIX.
 * Batched clouds always use DENSE (their packed space is compact); an
 * incremental update of a RADIX topology rebuilds it in full. */
enum gvdbx_topology_mode {
    GVDBX_TOPO_AUTO = 0,
    GVDBX_TOPO_DENSE = 1,
    GVDBX_TOPO_RADIX = 2
};

/* summary of the current topology, filled by gvdbx_get_info */
typedef struct {
    int levs;
    int depth;                 /* root level (DetermineDepth VG.cpp:821-837) */
    int top_lev;               /* lowest level with exactly one node (VG.cpp:3996-4012) */
    int root_pos[3];
    int64_t node_cnt[10];      /* nodes per level, root level included */
    int subdim[3];             /* atlas bricks per axis */
    int brick_res;             /* res0 + 2*apron */
    int apron;
    float pos_min[3], pos_max[3];   /* GetBoundingBox result (VG.cpp:5404-5467) */
    float obj_min[3], obj_max[3];   /* ComputeBounds result (VG.cpp:1792-1816) */
    int64_t sc_len;                 /* last InsertPointsSubcell list length */
    int topology_mode_used;         /* gvdbx_topology_mode actually taken */
    int log2dim[10];                /* Configure's log2 dimensions per level (0 beyond levs) */
} gvdbx_info;

/* ---- lifecycle ---- */
const char* gvdbx_last_error(void);
const char* gvdbx_version(void);
/* VolumeGVDB::VolumeGVDB + SetCudaDevice + Initialize (VG.cpp:240,2320-2361) */
int gvdbx_create(gvdbx_volume** out, int device);
int gvdbx_destroy(gvdbx_volume* v);
/* hipStream_t passed as void* so the ABI carries no HIP types */
int gvdbx_set_stream(gvdbx_volume* v, void* stream);
int gvdbx_synchronize(gvdbx_volume* v);

/* ---- configuration ---- */
/* Configure(levs, r, numcnt) VG.h:388 / VG.cpp:2380-2434; log2dims[0] is
 * the brick level.  Configure(q4,q3,q2,q1,q0) VG.h:387 == {q0,q1,q2,q3,q4}. */
int gvdbx_configure(gvdbx_volume* v, int levs, const int* log2dims);
/* SetChannelDefault VG.h:390 */
int gvdbx_set_channel_default(gvdbx_volume* v, int sx, int sy, int sz);
/* AddChannel VG.h:392 / VG.cpp:2437-2453 (GVDBX_FLOAT or GVDBX_UCHAR4;
 * filter/border are recorded for Render) */
int gvdbx_add_channel(gvdbx_volume* v, int chan, int dtype, int apron, int filter, int border);
/* FillChannel VG.cpp:2456-2470: float channels take value; uchar4 channels
 * take (value, value, value, 1) * 255 (gvdbOpFillC4,
 * cuda_gvdb_operators.cuh:368-373), or all zero when value == 0 */
int gvdbx_fill_channel(gvdbx_volume* v, int chan, float value);
/* choose the topology builder (no reference counterpart; default AUTO) */
int gvdbx_set_topology_mode(gvdbx_volume* v, int mode);
/* kernel tuning knobs (no reference counterpart): key "gather_variant"
 * (0 = LDS-staged candidates, 1 = ballot/readlane walk, 2 = per-lane
 * candidates with LDS minima; default 1), "gather_split" (variant 1: tiles
 * with more than twice this many 64-point candidate chunks are cut into
 * work items of about this many; 0 = one wavefront per tile; the result is
 * the same bit for bit; default 16, range [0, 2^20]), "gather_classes"
 * (variant 1: 1 = each bin's points grouped per call by the neighbour cells
 * they can reach, so a tile reads only the groups that can reach it; 0 =
 * every point of the 27 neighbour bins, whose tiles have about four times
 * the chunks: set gather_split to about 64 with it; same result bit for
 * bit; default 1),
 * "xcd_remap" (1 = contiguous brick
 * runs per XCD; default 0), "splat_variant" (SplatCubic: 0 = tile pull,
 * 1 = cell push; default 1; cell push needs a (res0 + 2 ceil(r))^3 float
 * grid within 64 KiB, else tile pull runs),
 * "splat_chunk" (cell push: points per wavefront run; the float summation
 * order, and so the last bits, depend on it; default 1024, range
 * [64, 2^24]), "sort_tiles" (the stable radix sort that bins points by
 * brick: 0 = tiles of 8192 entries from 2^23 valid points on, else 4096;
 * 1 = always 4096; 2 = always 8192; the bins are the same either way;
 * default 0), "bin_records" (1 = the binning sort of a single cloud moves
 * the 16-byte point records through its passes; 0 = it sorts (cell, index)
 * pairs and then gathers each point by index; the bins are the same either
 * way; default 1) */
int gvdbx_set_tuning(gvdbx_volume* v, const char* key, int value);
int gvdbx_get_tuning(gvdbx_volume* v, const char* key, int* value);

/* ---- points ---- */
/* SetPoints(DataPtr pos, vel, clr) VG.h:541 / VG.cpp:4900-4908 with
 * SetDataGPU VG.cpp:4872-4880: float3 positions at pos_dev + i*stride + off;
 * uchar4 colours at clr_dev + i*clr_stride + clr_off (4-byte aligned).
 * vel/clr may be NULL; velocities are not read by any stage of this path.
 * With colours, the gathers average them into their colour channel and
 * ScatterDensity writes them into channel 1 or 3, as the reference does. */
int gvdbx_set_points(gvdbx_volume* v, const void* pos_dev, int pos_off, int pos_stride,
                     const void* vel_dev, int vel_off, int vel_stride,
                     const void* clr_dev, int clr_off, int clr_stride);
/* per-point FP32 weights (extension; no reference counterpart) */
int gvdbx_set_point_weights(gvdbx_volume* v, const float* w_dev, int stride);

/* ---- topology ---- */
/* RebuildTopology(int, float, Vector3DF) VG.h:429 / VG.cpp:1181-1218.  A full
 * rebuild on every call, unless gvdbx_request_full_rebuild(v, 0) switched on
 * the reference's incremental branch (ActivateIncreBricksGPU VG.cpp:1020-1155
 * with bAccum = false): then, while the depth holds and every point lies
 * inside the root, leaves no point's box [int(p) - r, int(p) + r] reaches are
 * deactivated (mFlags 0, delinked, kept in the pool), missing nodes are
 * appended in key order; otherwise it rebuilds fully. */
int gvdbx_rebuild_topology(gvdbx_volume* v, int num_pnts, float radius, float ox, float oy, float oz);
/* AccumulateTopology(int, float, Vector3DF, int) VG.h:431 / VG.cpp:1158-1178:
 * the first call (or the first after gvdbx_request_full_rebuild(v, 1))
 * rebuilds; later calls add the nodes the points' boxes reach and keep every
 * existing one (ActivateIncreBricksGPU with bAccum = true).  The tree keeps
 * its depth and root: points outside the root give GVDBX_ERR_RANGE and no
 * change.  min_depth is accepted and has no further effect. */
int gvdbx_accumulate_topology(gvdbx_volume* v, int num_pnts, float radius, float ox, float oy, float oz,
                              int min_depth);
/* RequestFullRebuild(bool) VG.h:432: 1 (the default) makes every
 * RebuildTopology a full rebuild, 0 lets it update incrementally. */
int gvdbx_request_full_rebuild(gvdbx_volume* v, int tf);
/* FinishTopology(bool, bool) VG.h:405 / VG.cpp:1579-1593 */
int gvdbx_finish_topology(gvdbx_volume* v, int commit_pool, int compute_bounds);
/* UpdateAtlas() VG.h:406 / VG.cpp:2630-2704.  Active leaves take atlas slots
 * in leaf order (mValue, atlas map); leaves an incremental update
 * deactivated keep their earlier mValue and own no slot. */
int gvdbx_update_atlas(gvdbx_volume* v);

/* ---- points -> voxels ---- */
/* InsertPointsSubcell(int, int, float, Vector3DF, int&) VG.h:551 / VG.cpp:4971-5092 */
int gvdbx_insert_points_subcell(gvdbx_volume* v, int subcell_size, int num_pnts, float radius,
                                float tx, float ty, float tz, int64_t* out_len);
/* GatherDensity VG.h:556 / VG.cpp:5318-5344.  With point colours (SetPoints
 * clr), chan_clr must be a GVDBX_UCHAR4 channel and every interior voxel of
 * a subcell gets the mean colour of the subcell's list
 * (cuda_gvdb_particles.cuh:718-731): per component
 *   f = init + (float)S / 255,  stored: sat(trunc((f / (cnt + accum)) * 255))
 * where S is the exact integer sum over the list, init = (0,0,0,1) or the
 * stored colour / 255 when accumulating, and sat clamps to [0,255] (NaN -> 0;
 * an empty list without accumulate gives (0,0,0,255)).  Without point
 * colours chan_clr is not used. */
int gvdbx_gather_density(gvdbx_volume* v, int subcell_size, int num_pnts, float radius,
                         float tx, float ty, float tz, int64_t sc_len, int chan_density, int chan_clr, int accum);
/* GatherLevelSet VG.h:557 / VG.cpp:5348-5373 (colour as GatherDensity,
 * cuda_gvdb_particles.cuh:757-771) */
int gvdbx_gather_level_set(gvdbx_volume* v, int subcell_size, int num_pnts, float radius,
                           float tx, float ty, float tz, int64_t sc_len, int chan_levelset, int chan_clr, int accum);
/* GatherLevelSet that allocates nothing (no reference counterpart): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * the same interior voxels bit for bit as gvdbx_gather_level_set (which on
 * 8^3 bricks also fills the neighbours' aprons; this one leaves them to
 * gvdbx_update_apron, after which the bricks agree); its only scratch,
 * one 32-bit word per binned point, is a static device array of
 * GVDBX_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
 * the process ends: more than GVDBX_NOALLOC_SCRATCH_BYTES / 4 points (the
 * num_pnts of the last InsertPointsSubcell) fail with GVDBX_ERR_RANGE.  Calls on one GPU share that array, so calls that overlap
 * on one GPU (other threads or streams) race; calls ordered on one stream do
 * not.  Allocation happens only in the earlier calls that set the volume up
 * (topology, atlas, InsertPointsSubcell). */
#define GVDBX_NOALLOC_SCRATCH_BYTES ((uint64_t)64 << 20)
int gvdbx_gather_level_set_noalloc(gvdbx_volume* v, int subcell_size, int num_pnts, float radius,
                                   float tx, float ty, float tz, int64_t sc_len, int chan_levelset, int chan_clr,
                                   int accum);
/* InsertPointsSubcell_FP16 VG.h:552 / VG.cpp:5094-5189: the same lists
 * (and list length) as gvdbx_insert_points_subcell, with every listed
 * position kept as three 16-bit fractions of the bounding box of the last
 * RebuildTopology (GetBoundingBox's pos_min and pos_max - pos_min):
 *   q = ushort(((w - pos_min) / range) * 65535)   (truncated, clamped)
 * (gvdbInsertSubcell_fp16, cuda_gvdb_particles.cuh:154-216).  Membership is
 * decided from the full-precision position, as the reference inserts. */
int gvdbx_insert_points_subcell_fp16(gvdbx_volume* v, int subcell_size, int num_pnts, float radius,
                                     float tx, float ty, float tz, int64_t* out_len);
/* GatherLevelSet_FP16 VG.h:558 / VG.cpp:5375-5401
 * (gvdbGatherLevelSet_fp16, cuda_gvdb_particles.cuh:775-811): the level
 * set over FP16 lists, each position read back as
 * (q / 65535) * range + pos_min (FP32, in that order), starting from 3.0
 * (no accumulate).  Needs gvdbx_insert_points_subcell_fp16 and subcell
 * radius + subcell size <= 28; point colours are rejected (the reference
 * sums them into an uninitialised colour). */
int gvdbx_gather_level_set_fp16(gvdbx_volume* v, int subcell_size, int num_pnts, float radius,
                                float tx, float ty, float tz, int64_t sc_len, int chan_levelset, int chan_clr);
/* InsertPoints(int, Vector3DF, bool) VG.h:542 / VG.cpp:5219-5267 */
int gvdbx_insert_points(gvdbx_volume* v, int num_pnts, float tx, float ty, float tz, int prefix);
/* ScatterDensity(int, float, float, Vector3DF, bool, bool) VG.h:555 / VG.cpp:5624-5660.
 * Density goes to channel 0 (float).  With point colours, each point's home
 * voxel takes colour too: avg_color != 0 writes (Sx/n, Sy/n, Sz/n, 255)
 * (integer division of the voxel's uint32 sums) into channel 1
 * (gvdbScatterPointAvgCol, cuda_gvdb_particles.cuh:610-642); avg_color == 0
 * writes the colour of the voxel's highest-numbered point into channel 3
 * (cuda_gvdb_particles.cuh:570, whose race this resolves in point order).
 * That channel must be GVDBX_UCHAR4. */
int gvdbx_scatter_density(gvdbx_volume* v, int num_pnts, float radius, float amp,
                          float tx, float ty, float tz, int expand, int avg_color);
/* weighted cubic splat rho(v) = sum w_p max(0,1-|v-p|^2/r^2)^3 (extension of
 * the ScatterDensity kernel cuda_gvdb_particles.cuh:516-523) */
int gvdbx_splat_cubic(gvdbx_volume* v, int num_pnts, float radius, float tx, float ty, float tz,
                      int chan, int accum);
/* UpdateApron(uchar, float, bool) VG.h:408 / VG.cpp:4418-4457; on a uchar4
 * channel the value outside every leaf is make_uchar4(b, b, b, b) with b
 * converted by saturating truncation (cuda_gvdb_operators.cuh:139-142) */
int gvdbx_update_apron(gvdbx_volume* v, int chan, float boundval);

/* ---- autotuner (no reference counterpart) ---- */
/* Times the kernel configurations that leave the result unchanged on the
 * caller's own data, checks that each reproduces the configuration in
 * place bit for bit (a position-sensitive checksum of the output), keeps
 * the fastest and leaves it set.
 * Gather: GatherLevelSet (mode 1) or GatherDensity (mode 0) over the
 * current subcell lists into chan, across gather_variant {0, 1, 2},
 * gather_classes {0, 1}, gather_split {0, 8, 16, 32, 64, 128} and
 * xcd_remap {0, 1} (22 candidates); reps timed calls each.
 * Binning: InsertPointsSubcell(sc, num_pnts, radius) with the bins rebuilt
 * every call, across bin_records {0, 1} x sort_tiles {1, 2}; the bins, their
 * order and the list length must not change. */
typedef struct {
    int gather_variant, gather_classes, gather_split, xcd_remap;   /* chosen */
    float best_ms, default_ms;   /* per call: chosen, and the configuration in place before */
    int candidates;
} gvdbx_tune_gather;
typedef struct {
    int bin_records, sort_tiles;   /* chosen */
    float best_ms, default_ms;
    int candidates;
} gvdbx_tune_binning;
int gvdbx_autotune_gather(gvdbx_volume* v, int mode, int sc, float radius, int chan, int reps,
                          gvdbx_tune_gather* out);
int gvdbx_autotune_binning(gvdbx_volume* v, int num_pnts, int sc, float radius, int reps, gvdbx_tune_binning* out);

/* ---- frame graphs (no reference counterpart) ---- */
/* A hipGraph of one level-set frame on the current topology, for callers
 * that replay a fixed pipeline: depth frames accumulated into one topology,
 * fluid frames between topology rebuilds.  Create runs the frame once
 * (InsertPointsSubcell(sc, num_pnts, sc_radius), GatherLevelSet(radius,
 * chan, accum), UpdateApron(apron_chan, boundval) unless apron_chan < 0),
 * which sizes every buffer, then captures the same kernels (binning of the
 * points, subcell classes, gather, apron) on a stream of its own: the graph
 * holds no host step, allocation or memset.  Launch replays it on the
 * volume's stream over the current contents of the point buffer given to
 * SetPoints (same buffer, same count); it fails with GVDBX_ERR_STATE once
 * the topology, the point buffer or any device buffer changed since capture.
 * The list length of a replayed frame is not computed.  Point colours are
 * refused. */
typedef struct gvdbx_graph gvdbx_graph;
int gvdbx_frame_graph_create(gvdbx_volume* v, int num_pnts, int sc, float sc_radius, float radius, int chan,
                             int accum, int apron_chan, float boundval, gvdbx_graph** out);
int gvdbx_frame_graph_launch(gvdbx_graph* g);
int gvdbx_frame_graph_destroy(gvdbx_graph* g);

/* ---- files and transform ---- */
/* SaveVBX(std::string) VG.h:375 / VG.cpp:1626-1767: version 1.11, one float
 * grid in atlas layout with GVDB topology (GVDB_FILESPEC.txt, in the field
 * order SaveVBX writes: the channel count after the apron, the child-list
 * pool count per level).  Atlas bricks no active leaf owns are written as
 * zeros; the grid name is all zeros.  Channels must be 0..n-1. */
int gvdbx_save_vbx(gvdbx_volume* v, const char* path);
/* LoadVBX(std::string, int, int) VG.h:374 / VG.cpp:507-685: Configure from
 * the file, the node and child-list pools, FinishTopology, the channels
 * with the file's atlas size, UpdateAtlas; each active leaf takes the brick
 * at the atlas slot UpdateAtlas gives it.  Files with bitmask node records
 * (version 1.0, or 2.x with the flag) are refused; later grids replace
 * earlier ones, as there. */
int gvdbx_load_vbx(gvdbx_volume* v, const char* path);
/* LoadVBX(fname, force_maj, force_min) VG.h:374 / VG.cpp:507-554: as above,
 * reading the file as version force_maj.force_min (which decides whether a
 * transform and a bitmask flag follow the version bytes) when either is > 0,
 * whatever its header says. */
int gvdbx_load_vbx_version(gvdbx_volume* v, const char* path, int force_maj, int force_min);
/* SetTransform(pretrans, scale, angs, trans) VG.h:672: kept with the volume
 * and written to / read from VBX files; it changes nothing else here
 * (cameras and points are in index space). */
int gvdbx_set_transform(gvdbx_volume* v, const float* pretrans, const float* scale, const float* angs,
                        const float* trans);
int gvdbx_get_transform(gvdbx_volume* v, float* pretrans, float* scale, float* angs, float* trans);

/* ---- rendering (RC.cuh, DDA.cuh, GEO.cuh, MOD.cu) ---- */
enum gvdbx_shade {
    GVDBX_SHADE_VOXEL = 0,       /* gvdb_types.h:105 */
    GVDBX_SHADE_SECTION2D = 1,   /* gvdb_types.h:106 */
    GVDBX_SHADE_SECTION3D = 2,   /* gvdb_types.h:107 */
    GVDBX_SHADE_EMPTYSKIP = 3,   /* gvdb_types.h:108 */
    GVDBX_SHADE_TRILINEAR = 4,   /* gvdb_types.h:109 */
    GVDBX_SHADE_TRICUBIC = 5,    /* gvdb_types.h:110 */
    GVDBX_SHADE_LEVELSET = 6,    /* gvdb_types.h:111 */
    GVDBX_SHADE_VOLUME = 7,      /* gvdb_types.h:112 */
    GVDBX_SHADE_OFF = 100        /* gvdb_types.h:115 */
};

/* Camera in index space.  Pixel (x,y), u = (x+0.5)/width, w = (y+0.5)/height:
 *   origin    = (u*org_u + w*org_v) + org
 *   direction = normalize((u*dir_u + w*dir_v) + dir)
 * PrepareRender's perspective camera (VG.cpp:4260-4265; GEO.cuh:48-63) is
 * org = campos, org_u = org_v = 0, dir = cams, dir_u = camu, dir_v = camv;
 * an orthographic camera has dir_u = dir_v = 0. */
typedef struct {
    int width, height;
    float org[3], org_u[3], org_v[3];
    float dir[3], dir_u[3], dir_v[3];
} gvdbx_camera;

/* Scene parameters of ScnInfo (PrepareRender VG.cpp:4254-4306); defaults of
 * Scene::Scene (gvdb_scene.cpp:46-53) in brackets. */
typedef struct {
    float step;          /* SCN_DIRECTSTEP [1.0] */
    float extinct;       /* SCN_EXTINCT [-1.1]: a = exp(extinct * sigma * step) */
    float albedo;        /* SCN_ALBEDO [1.5] */
    float minval;        /* SCN_MINVAL [0.005] */
    float alphacut;      /* SCN_ALPHACUT [0.01] */
    float thresh, vmin, vmax;   /* SCN_THRESH, SCN_VMIN, SCN_VMAX [0.1, 0, 1] */
    float epsilon;       /* VDBInfo.epsilon [0.001] (VG.cpp:71) */
    int max_iter;        /* MAX_ITER [256] (RC.cuh:24) */
    float backclr[4];    /* SCN_BACKCLR */
    float light[3];      /* light position, index space */
    float shadow_amt, shadow_bias;   /* SCN_SHADOWAMT, SCN_SHADOWBIAS [0.8, 1.0] */
    const float* lut_dev;   /* transfer function: 16384 float4 in device memory (SHADE_VOLUME, sections) */
    float finestep;         /* SCN_FINESTEP [0.1] (SHADE_TRICUBIC) */
    float slice_pnt[3], slice_norm[3];   /* SCN_SLICE_PNT / SCN_SLICE_NORM (SetCrossSection, sections) */
} gvdbx_render_params;

/* ScnRay (cuda_gvdb_scene.cuh:66-74): the 64-byte ray record of Raytrace */
typedef struct {
    float hit[3], normal[3], orig[3], dir[3];
    uint32_t clr, pnode, pndx, pad;
} gvdbx_scn_ray;

/* Render(char shade, uchar chan, uchar buf) VG.h:341 / VG.cpp:4336-4381 with
 * ReadRenderBuf, every shading of the reference's switch: gvdbRayDeep
 * (SHADE_VOLUME), gvdbRaySurfaceVoxel / Trilinear / Tricubic, gvdbRayLevelSet,
 * gvdbRayEmptySkip, gvdbSection2D / 3D (MOD.cu:60-293) and SHADE_OFF (the
 * image cleared).  With a colour channel set (gvdbx_set_color_channel) the
 * surface shadings colour their hits, SHADE_VOLUME its samples and
 * SHADE_SECTION3D its slice and surface by that channel's uchar4 / 255 at
 * the voxel holding the point, as the reference does.  rgba_dev (optional): 4
 * floats per pixel, the colour before quantisation; rgba8_dev (optional):
 * the reference's uchar4 output (alpha 255 for EMPTYSKIP and the sections,
 * as there).  Rows of width pixels, y-major.  Tricubic taps past a padded
 * brick clamp to it (the reference reads into the neighbouring atlas brick). */
int gvdbx_render(gvdbx_volume* v, int shade, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp,
                 float* rgba_dev, uint8_t* rgba8_dev);
/* SetColorChannel(uchar chan) VG.h / VG.cpp:686-690: the uchar4 channel that
 * colours Render (getColorF RC.cuh:206-209); -1 or 255 (CHAN_UNDEF): none,
 * the default, and what loading a VBX file (DestroyChannels) resets it to.
 * Render fails with GVDBX_ERR_ARG while it names no uchar4 channel. */
int gvdbx_set_color_channel(gvdbx_volume* v, int chan);
/* mVDBInfo.clr_chan: the channel set above, -1 for none */
int gvdbx_get_color_channel(gvdbx_volume* v, int* chan);
/* Raytrace(DataPtr rays, uchar chan, char, int, float bias) VG.h:342 /
 * VG.cpp:4384-4408 (gvdbRaytrace MOD.cu:207-219): for every ray record in
 * device memory the trilinear surface hit (orig + t dir, first sample >=
 * thresh), pulled back by bias along dir, or (NOHIT, NOHIT, NOHIT); the
 * normal is written only on a hit. */
int gvdbx_raytrace(gvdbx_volume* v, int chan, const gvdbx_render_params* rp, gvdbx_scn_ray* rays_dev, int64_t num_rays,
                   float bias);
/* Emission-absorption raycast with the linear transfer s -> (s,s,s,s)
 * (extension for differentiation; same traversal and sample set as
 * SHADE_VOLUME, rayDeepBrick RC.cuh:485-533, no final clamp):
 *   a = exp(extinct*s*step); I += s*T*(1-a)*albedo; T *= a.
 * I_dev, T_dev: one float per pixel.  record != 0 keeps, per (ray, brick)
 * segment, the entry t, T and I for gvdbx_raycast_ea_backward. */
int gvdbx_raycast_ea(gvdbx_volume* v, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp,
                     float* I_dev, float* T_dev, int record);
/* Adjoint of the last recorded gvdbx_raycast_ea (same chan, cam, rp, I, T):
 * grad_dev (leaf-major padded bricks, nleaf * brick_res^3 floats) is
 * overwritten with d(sum gI*I + gT*T)/d voxel.  gT_dev may be NULL. */
int gvdbx_raycast_ea_backward(gvdbx_volume* v, int chan, const gvdbx_camera* cam, const gvdbx_render_params* rp,
                              const float* I_dev, const float* T_dev, const float* gI_dev, const float* gT_dev,
                              float* grad_dev);
/* Adjoint of UpdateApron: each apron gradient is added to the interior voxel
 * it copies (in place; apron entries are left as they are). */
int gvdbx_apron_adjoint(gvdbx_volume* v, float* grad_dev);
/* Adjoint of gvdbx_splat_cubic: from the interior voxel gradient grad_dev
 * (padded layout), dpos_dev (3 floats per point) and dw_dev (one per point,
 * may be NULL) for the points of the last SetPoints/SetPointWeights. */
int gvdbx_splat_cubic_backward(gvdbx_volume* v, int num_pnts, float radius, float tx, float ty, float tz,
                               const float* grad_dev, float* dpos_dev, float* dw_dev);

/* ---- depth frames -> points (extension for batched fusion) ---- */
/* Back-projects nframes depth images (row-major, width*height floats each,
 * metres) into index-space points, all frames in one launch.  intr4 = fx,
 * fy, cx, cy (host); poses_dev: 12 floats per frame, the camera-to-world
 * rotation R (row-major 3x3) then the translation t.  Pixel (u,v) with depth
 * d in [dmin, dmax] gives, in FP32 with no contraction,
 *   c = (d*((u-cx)/fx), d*((v-cy)/fy), d),
 *   w_i = ((R_i0*c.x + R_i1*c.y) + R_i2*c.z) + t_i,
 *   p_i = (w_i - origin_i) / voxel
 * (the world-to-index map of SetTransform, VG.h:672, with a uniform voxel
 * size); other pixels get the invalid point (0, 0, NOHIT), which every stage
 * skips.  out_xyz_dev: nframes*height*width float3. */
int gvdbx_backproject_depth(const float* depth_dev, int nframes, int width, int height, const float* intr4,
                            const float* poses_dev, float voxel, const float* origin3, float dmin, float dmax,
                            float* out_xyz_dev, void* stream);

/* ---- batched clouds (extension) ---- */
/* Many independent clouds in one volume: cloud c is the points
 * [cloud_off[c], cloud_off[c+1]) of pos_dev (host array of ncloud+1 offsets),
 * each in its own index space.  RebuildTopology then builds every cloud's
 * topology as the reference would for that cloud alone (its own bounding
 * box, depth, root and extra bricks) and packs the clouds into disjoint
 * slots of one index space, so every later stage runs once for all clouds.
 * The packing shifts cells by whole bricks (integers); positions, distances
 * and voxel values stay those of the cloud's own space.  SplatCubic,
 * UpdateApron, the adjoints and the batched raycasts work on batched clouds;
 * the subcell gathers, InsertPoints, ScatterDensity and Render do not. */
int gvdbx_set_points_batched(gvdbx_volume* v, const void* pos_dev, int pos_stride, const int64_t* cloud_off,
                             int ncloud);
/* per-cloud layout: cell_off (3 ints per cloud, level-0 cells added to the
 * cloud's cells in the packed space) and bounds (6 floats per cloud: its
 * ComputeBounds result in its own space); either may be NULL */
int gvdbx_read_cloud_layout(gvdbx_volume* v, int32_t* cell_off_host, float* bounds_host, int ncloud);
/* the cloud of every leaf (node_cnt[0] entries) */
int gvdbx_read_leaf_clouds(gvdbx_volume* v, int32_t* out_host, int64_t max_entries);
/* one camera per cloud (ncam == ncloud, equal image sizes), camera c in
 * cloud c's own space; image c at c*width*height of I_dev / T_dev */
int gvdbx_raycast_ea_batched(gvdbx_volume* v, int chan, const gvdbx_camera* cams, int ncam,
                             const gvdbx_render_params* rp, float* I_dev, float* T_dev, int record);
int gvdbx_raycast_ea_backward_batched(gvdbx_volume* v, int chan, const gvdbx_camera* cams, int ncam,
                                      const gvdbx_render_params* rp, const float* I_dev, const float* T_dev,
                                      const float* gI_dev, const float* gT_dev, float* grad_dev);

/* ---- readback (host copies; synchronising) ---- */
int gvdbx_get_info(gvdbx_volume* v, gvdbx_info* out);
/* pool group 0 of one level: 64-byte records (VG.h getNode / AL.cpp PoolFetch) */
int gvdbx_read_nodes(gvdbx_volume* v, int lev, void* out_host, int64_t max_nodes);
/* pool group 1 of one level: res^3 uint64 entries per node */
int gvdbx_read_child_lists(gvdbx_volume* v, int lev, uint64_t* out_host, int64_t max_entries);
/* atlas map, 16-byte entries, subdim.x*subdim.y*subdim.z of them */
int gvdbx_read_atlas_map(gvdbx_volume* v, void* out_host, int64_t max_entries);
/* all leaf bricks of a channel incl. apron, leaf-major, x fastest */
int gvdbx_read_bricks(gvdbx_volume* v, int chan, float* out_host, int64_t max_floats);
/* device pointer of a channel: brick of leaf k at k*brick_res^3 floats */
int gvdbx_channel_device_ptr(gvdbx_volume* v, int chan, float** out_dev);
/* point -> leaf index and rank within leaf from the last InsertPoints */
int gvdbx_read_point_nodes(gvdbx_volume* v, int32_t* pnode_host, int32_t* poff_host, int64_t n);

/* ---- host-side workload generators (no reference counterpart) ---- */
/* seeded synthetic inputs shared by tests and bench (SplitMix64 -> xoshiro256**) */
int gvdbx_gen_shell(uint64_t seed, int64_t n, float cx, float cy, float cz, float r0, float r1, float* out_xyz);
int gvdbx_gen_box(uint64_t seed, int64_t n, const float* lo, const float* hi, float jitter, float* out_xyz);

#ifdef __cplusplus
}
#endif
#endif
