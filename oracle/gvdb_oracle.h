/*
 * gvdb_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Single-threaded CPU restatement of the reference's point -> sparse-volume
 * path (GVDB Voxels 1.1.1).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product library (libgvdbx.so) never links or calls anything here.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference repository root:
 *   VG.cpp  = source/gvdb_library/src/gvdb_volume_gvdb.cpp
 *   AL.cpp  = source/gvdb_library/src/gvdb_allocator.cpp
 *   PT.cuh  = source/gvdb_library/kernels/cuda_gvdb_particles.cuh
 *   ND.cuh  = source/gvdb_library/kernels/cuda_gvdb_nodes.cuh
 *   OP.cuh  = source/gvdb_library/kernels/cuda_gvdb_operators.cuh
 *   RC.cuh  = source/gvdb_library/kernels/cuda_gvdb_raycast.cuh
 *   DDA.cuh = source/gvdb_library/kernels/cuda_gvdb_dda.cuh
 *   GEO.cuh = source/gvdb_library/kernels/cuda_gvdb_geom.cuh
 */
#ifndef GVDB_ORACLE_H
#define GVDB_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXLEV 10
#define OR_MAXCHAN 8
#define OR_ID_UNDEFL 0xFFFFFFFFu
#define OR_ID_UNDEF64 0xFFFFFFFFFFFFFFFFull
#define OR_NOHIT 1.0e10f

/* 64-byte node record, gvdb_node.h:27-40 == ND.cuh:24-35 */
typedef struct {
    uint8_t mLev, mFlags, mPriority, pad;
    int32_t mPos[3];
    int32_t mValue[3];
    float mVRange[3];
    uint64_t mParent;
    uint64_t mChildList;
    uint64_t mMask;
} or_node;

/* atlas map entry, ND.cuh:37-40 */
typedef struct {
    int32_t mPos[3];
    int32_t mLeafID;
} or_atlas_node;

typedef struct or_volume or_volume;

/* status codes (shared meaning with the product C-ABI) */
enum {
    OR_OK = 0,
    OR_ERR_ARG = 1,
    OR_ERR_STATE = 2,
    OR_ERR_DEPTH = 3,
    OR_ERR_NOPOINTS = 4,
    OR_ERR_RANGE = 5
};

or_volume* or_create(void);
void or_destroy(or_volume* v);

/* Configure(levs, r, numcnt) VG.cpp:2380-2434; logdim[0] is the brick level. */
int or_configure(or_volume* v, int levs, const int* logdim);
/* SetChannelDefault VG.h:390 */
void or_set_channel_default(or_volume* v, int sx, int sy, int sz);
/* channel types, gvdb_types.h:75-76 */
#define OR_T_UCHAR4 2
#define OR_T_FLOAT 3
/* AddChannel VG.cpp:2437-2453: or_add_channel makes a float channel */
int or_add_channel(or_volume* v, int chan, int apron);
int or_add_channel_typed(or_volume* v, int chan, int apron, int dtype);

/* RebuildTopology VG.cpp:1181-1218 (+ ActivateBricksGPU 1229-1413,
 * ActivateExtraBricksGPU 931-1016) followed by FinishTopology(false,true). */
int or_rebuild_topology(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes,
                        float radius, const float orig[3]);
/* ActivateIncreBricksGPU VG.cpp:1020-1155 on the current tree: accum = 1 is
 * AccumulateTopology (VG.cpp:1158-1178), accum = 0 the incremental
 * RebuildTopology (VG.cpp:1215-1216).  Keeps the tree's depth and root. */
int or_incremental_topology(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes, float radius,
                            const float orig[3], int accum);
/* UpdateAtlas VG.cpp:2630-2704 */
int or_update_atlas(or_volume* v);

/* InsertPointsSubcell VG.cpp:4971-5092; returns the list length in *out_len */
int or_insert_points_subcell(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride_bytes,
                             float radius, const float trans[3], int64_t* out_len);
/* InsertPointsSubcell_FP16 VG.cpp:5094-5189: positions stored as 16-bit
 * fractions of the bounding box (read back by or_gather: GatherLevelSet_FP16
 * is or_gather mode 1 without accumulate) */
int or_insert_points_subcell_fp16(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride_bytes,
                                  float radius, const float trans[3], int64_t* out_len);
/* GatherDensity VG.cpp:5318-5344 / PT.cuh:696-733 (mode 0)
 * GatherLevelSet VG.cpp:5348-5373 / PT.cuh:736-773 (mode 1) */
int or_gather(or_volume* v, int mode, float radius, int chan, int accum);
/* same semantics evaluated in double; writes one double per interior leaf
 * voxel, brick-major, x fastest (leaf k, voxel (x,y,z) at k*res^3 + (z*res+y)*res+x) */
int or_gather_f64(or_volume* v, int mode, float radius, int accum, double* out);

/* UpdateApron VG.cpp:4418-4457 / OP.cuh:73-126 */
int or_update_apron(or_volume* v, int chan, float boundval);

/* InsertPoints VG.cpp:5219-5267 / PT.cuh:19-36 (poff = stable rank) */
int or_insert_points(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes,
                     const float trans[3], int32_t* pnode, int32_t* poff);
/* ScatterDensity VG.cpp:5624-5660 / PT.cuh:516-573 (serial sum) */
int or_scatter_density(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes,
                       float radius, float amp, const float trans[3], int expand, int chan);
/* weighted cubic splat (extension of PT.cuh:516-523 distFunc) */
int or_splat_cubic(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes,
                   const float* weights, float radius, const float trans[3], int chan, int accum);
int or_splat_cubic_f64(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes,
                       const float* weights, float radius, const float trans[3], double* out);
/* FillChannel VG.cpp:2456-2470 (float; uchar4: zero, or (val,val,val,1)*255) */
int or_fill_channel(or_volume* v, int chan, float val);
/* colour mean of the gathers (PT.cuh:718-731,757-771, defined per D8 in
 * gvdb_oracle.c) over the lists of the last or_insert_points_subcell;
 * clr: uchar4 per point at clr + i*clr_stride */
int or_gather_color(or_volume* v, const uint8_t* clr, int64_t clr_stride, int chan, int accum);
/* colour part of ScatterDensity (PT.cuh:554-572,610-642; D9) */
int or_scatter_color(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes, const uint8_t* clr,
                     int64_t clr_stride, const float trans[3], int avg, int chan);

/* ---- readback ---- */
typedef struct {
    int levs;
    int depth;
    int top_lev;
    int root_pos[3];
    int64_t node_cnt[OR_MAXLEV];
    int subdim[3];
    int brick_res;   /* res0 + 2*apron */
    int apron;
    float pos_min[3], pos_max[3];   /* GetBoundingBox result */
    float obj_min[3], obj_max[3];   /* ComputeBounds result  */
    int64_t sc_len;
} or_info;

int or_get_info(or_volume* v, or_info* out);
const or_node* or_nodes(or_volume* v, int lev);
const uint64_t* or_child_lists(or_volume* v, int lev);
const or_atlas_node* or_atlas_map(or_volume* v, int64_t* count);
/* copy leaf k's full brick (res0+2a)^3, x fastest, into out */
/* SaveVBX VG.cpp:1626-1767 (version 1.11, one float grid, atlas layout) */
int or_save_vbx(or_volume* v, const char* path);
/* test input: overwrite one leaf's padded brick (inverse of or_read_brick) */
int or_write_brick(or_volume* v, int chan, int64_t leaf, const float* in);
int or_read_brick(or_volume* v, int chan, int64_t leaf, float* out);
/* pointer to the full 3-D atlas, dims subdim*(res0+2a), x fastest */
const float* or_atlas(or_volume* v, int chan);

/* ---- raycast (RC.cuh, DDA.cuh, GEO.cuh, MOD.cu) ---- */
#define OR_SHADE_VOXEL 0       /* gvdb_types.h:105 */
#define OR_SHADE_SECTION2D 1   /* gvdb_types.h:106 */
#define OR_SHADE_SECTION3D 2   /* gvdb_types.h:107 */
#define OR_SHADE_EMPTYSKIP 3   /* gvdb_types.h:108 */
#define OR_SHADE_TRILINEAR 4   /* gvdb_types.h:109 */
#define OR_SHADE_TRICUBIC 5    /* gvdb_types.h:110 */
#define OR_SHADE_LEVELSET 6    /* gvdb_types.h:111 */
#define OR_SHADE_VOLUME 7      /* gvdb_types.h:112 */
#define OR_SHADE_OFF 100       /* gvdb_types.h:115 */

/* Camera: the ray of pixel (x,y), u = (x+0.5)/width, w = (y+0.5)/height, has
 *   origin    = (u*org_u + w*org_v) + org
 *   direction = normalize((u*dir_u + w*dir_v) + dir)
 * in index space.  A perspective camera (GEO.cuh:48-63: campos, cams, camu,
 * camv) is org = campos, org_u = org_v = 0, dir = cams, dir_u = camu,
 * dir_v = camv; an orthographic one has dir_u = dir_v = 0. */
typedef struct {
    int width, height;
    float org[3], org_u[3], org_v[3];
    float dir[3], dir_u[3], dir_v[3];
} or_camera;

/* Scene parameters (ScnInfo, PrepareRender VG.cpp:4254-4306) */
typedef struct {
    int shade;          /* OR_SHADE_* (or_render only) */
    float step;         /* SCN_DIRECTSTEP */
    float extinct;      /* SCN_EXTINCT (negative: a = exp(extinct*sigma*step)) */
    float albedo;       /* SCN_ALBEDO */
    float minval;       /* SCN_MINVAL */
    float alphacut;     /* SCN_ALPHACUT */
    float thresh, vmin, vmax;   /* SCN_THRESH, SCN_VMIN, SCN_VMAX */
    float epsilon;      /* VDBInfo.epsilon (VG.cpp:71) */
    int max_iter;       /* MAX_ITER (RC.cuh:24) */
    float backclr[4];   /* SCN_BACKCLR */
    float light[3];     /* light position in index space */
    float shadow_amt, shadow_bias;   /* SCN_SHADOWAMT, SCN_SHADOWBIAS */
    const float* lut;   /* transfer function, 16384 x float4 (SHADE_VOLUME, sections) */
    float finestep;     /* SCN_FINESTEP [0.1] (SHADE_TRICUBIC) */
    float slice_pnt[3], slice_norm[3];   /* SCN_SLICE_PNT / SCN_SLICE_NORM (sections) */
    int use_clr, clr_chan;   /* SetColorChannel VG.cpp:686-690: use_clr != 0 colours or_render from uchar4 channel clr_chan */
} or_render_params;

/* ScnRay (cuda_gvdb_scene.cuh:66-74): 64-byte ray record of Raytrace */
typedef struct {
    float hit[3], normal[3], orig[3], dir[3];
    uint32_t clr, pnode, pndx, pad;
} or_scn_ray;

/* Render(shade) VG.cpp:4336-4381: gvdbRayDeep / gvdbRaySurfaceTrilinear /
 * gvdbRayLevelSet (MOD.cu:60-181).  out32: 4 floats per pixel, the colour
 * before the uchar4 quantisation.  out64 (optional): the same pixel evaluated
 * in FP64 on the FP32-defined sample set (sample positions, transfer-function
 * indices, hits and shadow decisions).  bricks64 (optional): FP64 padded
 * bricks to sample instead of the FP32 atlas (leaf-major, x fastest). */
int or_render(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp, const double* bricks64,
              float* out32, double* out64);

/* Raytrace VG.cpp:4384-4408 / gvdbRaytrace MOD.cu:207-219: every ray's hit
 * with the trilinear surface brick, pulled back by bias along the ray */
int or_raytrace(or_volume* v, int chan, const or_render_params* rp, or_scn_ray* rays, int64_t n, float bias);

/* Emission-absorption raycast with the linear transfer s -> (s,s,s,s)
 * (extension; same traversal and samples as SHADE_VOLUME, no final clamp):
 *   a = exp(extinct*s*step); I += s*T*(1-a)*albedo; T *= a.
 * One I and one T per pixel, FP32 and FP64 (any output may be NULL). */
int or_raycast_ea(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp, const double* bricks64,
                  float* I32, float* T32, double* I64, double* T64);
/* FP64 adjoint: grad (nb*br^3 padded bricks) += d(sum gI*I + gT*T)/d voxel */
int or_raycast_ea_backward(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp,
                           const double* bricks64, const double* gI, const double* gT, double* grad);
/* magnitude form of the adjoint: every term replaced by the magnitude of its
 * operands (|gI|, |gT|, |1 - a| + |e s a|, |e| times the magnitude of the
 * I sum), so that
 * n * 2^-24 * result bounds the rounding error of an FP32 evaluation whose
 * longest chain of dependent operations has n steps */
int or_raycast_ea_backward_mag(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp,
                               const double* bricks64, const double* gI, const double* gT, double* grad);
/* the samples of one pixel's ray: leaf ids and brick-local positions */
int or_raycast_ea_samples(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp, int64_t px,
                          int64_t max_out, int64_t* leaf_out, float* p_out, int64_t* count);

/* FP64 padded-brick chain used by the adjoint tests */
int or_apron_f64(or_volume* v, double* bricks, double boundval);
int or_apron_adjoint_f64(or_volume* v, double* grad);
int or_splat_cubic_f64_padded(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                              double* bricks);
int or_splat_cubic_backward_f64(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                                const double* grad, double* dpos, double* dw);
/* magnitude form: sums of |g| |w| |dk/dp| and |g| k */
int or_splat_cubic_backward_mag_f64(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                                    const double* grad, double* dpos, double* dw);

#ifdef __cplusplus
}
#endif
#endif
