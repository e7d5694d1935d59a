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
/* AddChannel VG.cpp:2437-2453 (float channels only) */
int or_add_channel(or_volume* v, int chan, int apron);

/* RebuildTopology VG.cpp:1181-1218 (+ ActivateBricksGPU 1229-1413,
 * ActivateExtraBricksGPU 931-1016) followed by FinishTopology(false,true). */
int or_rebuild_topology(or_volume* v, const float* pos, int64_t n, int64_t stride_bytes,
                        float radius, const float orig[3]);
/* UpdateAtlas VG.cpp:2630-2704 */
int or_update_atlas(or_volume* v);

/* InsertPointsSubcell VG.cpp:4971-5092; returns the list length in *out_len */
int or_insert_points_subcell(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride_bytes,
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
/* FillChannel (float) */
int or_fill_channel(or_volume* v, int chan, float val);

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
int or_read_brick(or_volume* v, int chan, int64_t leaf, float* out);
/* pointer to the full 3-D atlas, dims subdim*(res0+2a), x fastest */
const float* or_atlas(or_volume* v, int chan);

/* ---- raycast (emission-absorption / trilinear) ---- */
typedef struct {
    float extinct;    /* SCN_EXTINCT (negative in the reference) */
    float albedo;     /* SCN_ALBEDO */
    float step;       /* SCN_DIRECTSTEP */
    float minval;     /* SCN_MINVAL */
    float alphacut;   /* SCN_ALPHACUT */
    float thresh, vmin, vmax;  /* SCN_THRESH/VMIN/VMAX */
    float epsilon;
    int max_iter;
    float backclr[4];
} or_render_params;

/* Orthographic emission-absorption raycast with analytic transfer
 * (sigma = -extinct * density, emission = density). rays given as
 * origin/direction arrays (3 floats each).  Output rgba (4 floats per ray).
 * Follows rayCast RC.cuh:543-611 / rayDeepBrick RC.cuh:485-533 with
 * software trilinear sampling. */
int or_raycast_ea(or_volume* v, int chan, const or_render_params* rp, int64_t nrays,
                  const float* origins, const float* dirs, float* out_rgba, double* out_f64);

#ifdef __cplusplus
}
#endif
#endif
