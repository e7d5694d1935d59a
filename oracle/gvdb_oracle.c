/*
 * gvdb_oracle.c -- TEST INFRASTRUCTURE ONLY (see gvdb_oracle.h).
 *
 * A single-threaded C restatement of the reference's point -> sparse-volume
 * path.  It follows the reference's own algorithm (sort keys, unique, pool
 * allocation in key order, top-down linking by tree descent, subcell point
 * lists, per-subcell gather), not the GPU design of the product, so that the
 * two are independent implementations of one semantics.
 *
 * Floating-point conventions (fixed here and in the product kernels):
 *   - every expression is evaluated in IEEE binary32 with round-to-nearest and
 *     no contraction (built with -ffp-contract=off);
 *   - a squared distance is fmaf(dz,dz, fmaf(dy,dy, dx*dx));
 *   - sqrtf and divisions are IEEE correctly rounded.
 * The reference builds with --use_fast_math, whose results are not
 * reproducible off NVIDIA hardware; float results are therefore pinned to
 * these defined semantics and, separately, to an FP64 evaluation.
 *
 * Deliberate deviations (the reference is undefined or racy there):
 *   D1 invalid points (non-finite, z == NOHIT before translation, or any
 *      coordinate < 0 after translation) are dropped from every stage,
 *      including the bounding box (reference: VG.cpp:5404-5467 includes them);
 *   D2 a child-list read outside the allocated lists returns "empty"
 *      (reference reads stale pool memory, ND.cuh:115-124);
 *   D3 the extra-brick existence test is "the node is in the phase-1 node
 *      set" (reference getPosParent, ND.cuh:303-348, which equals this when
 *      the descent is defined) and extra nodes outside the root node are not
 *      emitted (reference: out-of-range child-list reads);
 *   D4 a point -> leaf lookup that lands on a leaf not containing the
 *      position is "no leaf" (reference: out-of-range subcell index writes,
 *      PT.cuh:255-266, or unchecked root when top_lev==0, ND.cuh:199-226);
 *   D5 list order and poff are by ascending point index (reference: atomic
 *      order, PT.cuh:35,266);
 *   D6 ScatterDensity accumulates serially in point order (reference: racy
 *      surface read-modify-write, PT.cuh:543-551).
 *   D8 colour means (GatherDensity / GatherLevelSet with point colours,
 *      PT.cuh:718-731,757-771) sum the listed points' uchar4 components as
 *      exact integers and convert once: f = init + (float)S / 255, mean =
 *      f / (float)(count + accumulate), stored as a saturating truncation of
 *      mean * 255 (NaN -> 0); an empty list without accumulate gives
 *      (0,0,0,255), which is what 0/0 and 1/0 saturate to (reference: an
 *      FP32 sum of c/255 in atomic list order, then the same division);
 *   D9 ScatterDensity without colour averaging writes the colour of the
 *      highest-numbered point of each voxel (reference: racy surface write,
 *      PT.cuh:570);
 *   D10 the incremental update keeps the tree's depth and root: points
 *      outside the root are refused with no change (reference: out-of-range
 *      child-list reads), box cells outside the root are not emitted (D3);
 *   D11 tricubic taps are formed in brick-local coordinates and clamp to the
 *      padded brick (reference: floor(p + atlas offset) in FP32, reading on
 *      into the neighbouring atlas brick).
 */
#include "gvdb_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct or_volume {
    int levs;
    int logdim[OR_MAXLEV];
    int res[OR_MAXLEV];
    int64_t range[OR_MAXLEV];
    /* pools: group 0 nodes, group 1 child lists (AL.cpp:55-231) */
    or_node* nodes[OR_MAXLEV];
    int64_t ncnt[OR_MAXLEV], ncap[OR_MAXLEV];
    uint64_t* clist[OR_MAXLEV];
    int64_t ccnt[OR_MAXLEV], ccap[OR_MAXLEV];
    int depth, top_lev;
    int root_pos[3];
    float pos_min[3], pos_max[3];
    float obj_min[3], obj_max[3];
    int configured;
    /* atlas (AL.cpp:528-715) */
    int default_axis[3];
    int subdim[3];
    int apron;
    int has_chan[OR_MAXCHAN];
    float* atlas[OR_MAXCHAN];       /* 4-byte words: float, or packed uchar4 (x lowest byte) */
    int chan_type[OR_MAXCHAN];      /* OR_T_FLOAT or OR_T_UCHAR4 */
    int64_t atlas_vox;
    or_atlas_node* amap;
    int64_t amap_cnt;
    int64_t atlas_used;
    /* subcell lists (VG.cpp:4971-5092) */
    int sc_size, sc_per_brick, sc_dim;
    int64_t num_sc;
    int64_t sc_len;
    int32_t* sc_cnt;
    int64_t* sc_off;
    float* sc_pnt;
    int64_t* sc_pidx;   /* point index of each list entry (colour lookup) */
    int32_t* sc_nid;
    int32_t* sc_pos;
};

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */

static int64_t ipow3(int64_t r) { return r * r * r; }

/* float -> uchar as CUDA's saturating conversion: truncation toward zero,
 * clamped to [0, 255], NaN -> 0 */
static uint32_t sat_u8(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 255.0f) return 255;
    return (uint32_t)x;
}

/* the word UpdateApron stores outside every leaf: the float itself, or
 * make_uchar4(b, b, b, b) (OP.cuh:139-142) */
static uint32_t bound_word(int dtype, float boundval) {
    uint32_t w;
    if (dtype == OR_T_UCHAR4) {
        uint32_t c = sat_u8(boundval);
        w = c | c << 8 | c << 16 | c << 24;
    } else {
        memcpy(&w, &boundval, 4);
    }
    return w;
}

/* getRange VG.h:614-619 */
static int64_t or_range(const or_volume* v, int lev) {
    if (lev < 0) return 1;
    int64_t r = v->res[0];
    for (int l = 1; l <= lev; l++) r *= v->res[l];
    return r;
}

/* D1: validity of a raw input point under translation */
static int point_valid(const float* p, const float t[3], float w[3]) {
    if (!isfinite(p[0]) || !isfinite(p[1]) || !isfinite(p[2])) return 0;
    if (p[2] == OR_NOHIT) return 0;
    w[0] = p[0] + t[0];
    w[1] = p[1] + t[1];
    w[2] = p[2] + t[2];
    if (!isfinite(w[0]) || !isfinite(w[1]) || !isfinite(w[2])) return 0;
    if (w[0] < 0 || w[1] < 0 || w[2] < 0) return 0;
    return 1;
}

static const float* pt(const float* base, int64_t i, int64_t stride) {
    return (const float*)((const char*)base + i * stride);
}

/* PoolAlloc AL.cpp:163-192 (capacity doubling; contents zeroed) */
static int64_t pool_alloc_node(or_volume* v, int lev) {
    if (v->ncnt[lev] >= v->ncap[lev]) {
        int64_t nc = v->ncap[lev] ? v->ncap[lev] * 2 : 4;
        or_node* nn = (or_node*)calloc((size_t)nc, sizeof(or_node));
        if (v->nodes[lev]) {
            memcpy(nn, v->nodes[lev], (size_t)v->ncnt[lev] * sizeof(or_node));
            free(v->nodes[lev]);
        }
        v->nodes[lev] = nn;
        v->ncap[lev] = nc;
    }
    int64_t i = v->ncnt[lev]++;
    memset(&v->nodes[lev][i], 0, sizeof(or_node));
    return i;
}

static int64_t pool_alloc_clist(or_volume* v, int lev) {
    int64_t len = ipow3(v->res[lev]);
    if (v->ccnt[lev] >= v->ccap[lev]) {
        int64_t nc = v->ccap[lev] ? v->ccap[lev] * 2 : 4;
        uint64_t* nn = (uint64_t*)malloc((size_t)(nc * len) * sizeof(uint64_t));
        if (v->clist[lev]) {
            memcpy(nn, v->clist[lev], (size_t)(v->ccnt[lev] * len) * sizeof(uint64_t));
            free(v->clist[lev]);
        }
        v->clist[lev] = nn;
        v->ccap[lev] = nc;
    }
    int64_t i = v->ccnt[lev]++;
    /* ActivateBricksGPU fills every new child list with 0xFF (VG.cpp:1327-1329) */
    memset(v->clist[lev] + i * len, 0xFF, (size_t)len * sizeof(uint64_t));
    return i;
}

static uint64_t elem(int grp, int lev, int64_t ndx) {
    /* Elem AL.h:59-62 */
    return (uint64_t)grp | ((uint64_t)lev << 8) | ((uint64_t)ndx << 16);
}

/* SetupNode VG.cpp:2507-2520 */
static void setup_node(or_volume* v, int lev, int64_t idx, int px, int py, int pz) {
    or_node* n = &v->nodes[lev][idx];
    n->mLev = (uint8_t)lev;
    n->mPos[0] = px; n->mPos[1] = py; n->mPos[2] = pz;
    n->mChildList = OR_ID_UNDEFL;
    n->mParent = OR_ID_UNDEFL;
    n->mValue[0] = n->mValue[1] = n->mValue[2] = -1;
    n->mFlags = 1;
}

/* getChild ND.cuh:115-124, with D2 for reads outside the allocated lists */
static int64_t get_child(const or_volume* v, const or_node* nd, int64_t b) {
    uint64_t listid = nd->mChildList;
    if (listid == OR_ID_UNDEFL) return -1;
    int clev = (int)((listid >> 8) & 0xFF);
    int64_t cndx = (int64_t)(listid >> 16);
    int64_t len = ipow3(v->res[clev]);
    int64_t flat = cndx * len + b;
    if (clev <= 0 || clev >= v->levs) return -1;
    if (flat < 0 || flat >= v->ccnt[clev] * len) return -1;
    uint64_t e = v->clist[clev][flat];
    if (e == OR_ID_UNDEF64) return -1;
    return (int64_t)(int32_t)(uint32_t)(e >> 16);
}

/* integer (pos - npos)*res/range with C truncation, ND.cuh:318-323 */
static void pos_in_node(const int pos[3], const int32_t np[3], int res, int64_t range, int64_t out[3]) {
    for (int a = 0; a < 3; a++) {
        int64_t d = (int64_t)pos[a] - np[a];
        d *= res;
        out[a] = d / range;
    }
}

/* getParent ND.cuh:350-397 (used by LinkBricks) */
static int64_t get_parent(const or_volume* v, int stop_lev, const int pos[3]) {
    if (v->ncnt[v->top_lev] < 1) return -1;
    const or_node* nd = &v->nodes[v->top_lev][0];
    if (!nd->mFlags) return 0;
    int l = nd->mLev;
    if (l == stop_lev) return 0;
    while (l > stop_lev) {
        int res = v->res[l];
        int64_t pin[3];
        pos_in_node(pos, nd->mPos, res, v->range[l], pin);
        int64_t bit = (pin[2] * res + pin[1]) * res + pin[0];
        l--;
        int64_t c = get_child(v, nd, bit);
        if (c < 0) return -1;
        if (l == stop_lev) return c;
        if (c >= v->ncnt[l]) return -1;
        nd = &v->nodes[l][c];
    }
    return -1;
}

/* top level selection of PrepareVDBPartially VG.cpp:3996-4012 */
static void compute_top_lev(or_volume* v) {
    int tlev = 1;
    for (int n = v->levs - 1; n >= 0; n--)
        if (v->ncnt[n] == 1) tlev = n;
    v->top_lev = tlev;
}

/* gvdbLinkBricks ND.cuh:470-505, launched top-down (VG.cpp:1394-1402) */
static void link_level(or_volume* v, int lev) {
    if (lev + 1 >= v->levs) return;
    int res = v->res[lev + 1];
    int64_t range = v->range[lev + 1];
    int64_t len = ipow3(res);
    for (int64_t i = 0; i < v->ncnt[lev]; i++) {
        or_node* node = &v->nodes[lev][i];
        if (!node->mFlags) continue;
        int pos[3] = {node->mPos[0], node->mPos[1], node->mPos[2]};
        int64_t p = get_parent(v, lev + 1, pos);
        if (p < 0 || p >= v->ncnt[lev + 1]) continue;
        or_node* pn = &v->nodes[lev + 1][p];
        int64_t pin[3];
        pos_in_node(pos, pn->mPos, res, range, pin);
        int64_t bit = (pin[2] * res + pin[1]) * res + pin[0];
        if (pin[0] > res || pin[0] < 0 || pin[1] > res || pin[1] < 0 || pin[2] > res || pin[2] < 0) continue;
        node->mParent = ((uint64_t)p << 16) | ((uint64_t)(lev + 1) << 8);
        int64_t cndx = (int64_t)(pn->mChildList >> 16);
        if (cndx >= v->ncnt[lev + 1]) continue;
        int64_t flat = cndx * len + bit;
        if (flat < 0 || flat >= v->ccnt[lev + 1] * len) continue; /* D2 */
        v->clist[lev + 1][flat] = ((uint64_t)i << 16) | ((uint64_t)lev << 8);
    }
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

/* GetCoveringNode (float version) PT.cuh:88-100 */
static int cover_axis(float p, int range) {
    float q = p / (float)range;
    int np = (int)(ceilf(q) * (float)range);
    if (p < (float)np) np -= range;
    return np;
}

static uint64_t make_key(int lev, int kx, int ky, int kz) {
    /* u16 quad {z,y,x,lev} read as little-endian u64 (PT.cuh:377-380, VG.cpp:916-923) */
    return ((uint64_t)(uint16_t)lev << 48) | ((uint64_t)(uint16_t)kx << 32) |
           ((uint64_t)(uint16_t)ky << 16) | (uint64_t)(uint16_t)kz;
}

static int64_t unique_u64(uint64_t* a, int64_t n) {
    /* gvdbFindUnique + gvdbCompactUnique PT.cuh:336-350,934-943 */
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (i == 0 || a[i] != a[i - 1]) a[m++] = a[i];
    return m;
}

/* allocate nodes of one sorted-unique key set, per level in key order
 * (ActivateBricksGPU VG.cpp:1305-1331, ActivateExtraBricksGPU 966-988) */
static void alloc_from_keys(or_volume* v, const uint64_t* keys, int64_t m) {
    for (int lev = 0; lev < v->depth; lev++) {
        int64_t rng = v->range[lev];
        for (int64_t j = 0; j < m; j++) {
            uint64_t k = keys[j];
            if ((int)((k >> 48) & 0xFFFF) != lev) continue;
            int kx = (int)((k >> 32) & 0xFFFF), ky = (int)((k >> 16) & 0xFFFF), kz = (int)(k & 0xFFFF);
            int64_t id = pool_alloc_node(v, lev);
            setup_node(v, lev, id, (int)(kx * rng), (int)(ky * rng), (int)(kz * rng));
            if (lev > 0) {
                int64_t c = pool_alloc_clist(v, lev);
                v->nodes[lev][id].mChildList = elem(1, lev, c);
            }
        }
    }
}

static int key_in(const uint64_t* keys, int64_t m, uint64_t k) {
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (keys[mid] < k) lo = mid + 1; else hi = mid;
    }
    return lo < m && keys[lo] == k;
}

/* ------------------------------------------------------------------ */
/* lifecycle / configuration                                           */
/* ------------------------------------------------------------------ */

or_volume* or_create(void) {
    or_volume* v = (or_volume*)calloc(1, sizeof(or_volume));
    v->default_axis[0] = 16; v->default_axis[1] = 16; v->default_axis[2] = 1;  /* VG.cpp:2348 */
    v->top_lev = 1;
    return v;
}

static void free_sc(or_volume* v) {
    free(v->sc_cnt); free(v->sc_off); free(v->sc_pnt); free(v->sc_pidx); free(v->sc_nid); free(v->sc_pos);
    v->sc_cnt = NULL; v->sc_off = NULL; v->sc_pnt = NULL; v->sc_pidx = NULL; v->sc_nid = NULL; v->sc_pos = NULL;
    v->num_sc = 0; v->sc_len = 0;
}

void or_destroy(or_volume* v) {
    if (!v) return;
    for (int l = 0; l < OR_MAXLEV; l++) { free(v->nodes[l]); free(v->clist[l]); }
    for (int c = 0; c < OR_MAXCHAN; c++) free(v->atlas[c]);
    free(v->amap);
    free_sc(v);
    free(v);
}

int or_configure(or_volume* v, int levs, const int* logdim) {
    if (!v || levs < 2 || levs > OR_MAXLEV || !logdim) return OR_ERR_ARG;
    for (int l = 0; l < OR_MAXLEV; l++) {
        free(v->nodes[l]); free(v->clist[l]);
        v->nodes[l] = NULL; v->clist[l] = NULL;
        v->ncnt[l] = v->ncap[l] = v->ccnt[l] = v->ccap[l] = 0;
    }
    v->levs = levs;
    for (int l = 0; l < levs; l++) {
        v->logdim[l] = logdim[l] == 0 ? 1 : logdim[l];
        if (v->logdim[l] < 1 || v->logdim[l] > 10) return OR_ERR_ARG;
        v->res[l] = 1 << v->logdim[l];
    }
    for (int l = 0; l < levs; l++) v->range[l] = or_range(v, l);
    v->configured = 1;
    v->depth = 0;
    v->top_lev = 1;
    return OR_OK;
}

void or_set_channel_default(or_volume* v, int sx, int sy, int sz) {
    v->default_axis[0] = sx; v->default_axis[1] = sy; v->default_axis[2] = sz;
}

static int brick_res(const or_volume* v) { return v->res[0] + 2 * v->apron; }

static void atlas_resize(or_volume* v, int sz) {
    /* AtlasResize AL.cpp:622-650: grows Z, preserves the old contents */
    int br = brick_res(v);
    int64_t nvox = (int64_t)v->subdim[0] * br * (int64_t)v->subdim[1] * br * (int64_t)sz * br;
    for (int c = 0; c < OR_MAXCHAN; c++) {
        if (!v->has_chan[c]) continue;
        float* na = (float*)calloc((size_t)nvox, sizeof(float));
        if (v->atlas[c]) {
            int64_t keep = v->atlas_vox < nvox ? v->atlas_vox : nvox;
            memcpy(na, v->atlas[c], (size_t)keep * sizeof(float));
            free(v->atlas[c]);
        }
        v->atlas[c] = na;
    }
    v->subdim[2] = sz;
    v->atlas_vox = nvox;
}

int or_add_channel(or_volume* v, int chan, int apron) { return or_add_channel_typed(v, chan, apron, OR_T_FLOAT); }

int or_add_channel_typed(or_volume* v, int chan, int apron, int dtype) {
    if (!v || !v->configured || chan < 0 || chan >= OR_MAXCHAN || apron < 0) return OR_ERR_ARG;
    if (dtype != OR_T_FLOAT && dtype != OR_T_UCHAR4) return OR_ERR_ARG;
    int any = 0;
    for (int c = 0; c < OR_MAXCHAN; c++) any |= v->has_chan[c];
    if (!any) {
        v->subdim[0] = v->default_axis[0]; v->subdim[1] = v->default_axis[1]; v->subdim[2] = v->default_axis[2];
        v->apron = apron;
        v->atlas_vox = 0;
    } else if (apron != v->apron) {
        return OR_ERR_ARG;
    }
    v->has_chan[chan] = 1;
    v->chan_type[chan] = dtype;
    int br = brick_res(v);
    int64_t nvox = (int64_t)v->subdim[0] * br * (int64_t)v->subdim[1] * br * (int64_t)v->subdim[2] * br;
    free(v->atlas[chan]);
    v->atlas[chan] = (float*)calloc((size_t)nvox, sizeof(float));
    v->atlas_vox = nvox;
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* topology                                                            */
/* ------------------------------------------------------------------ */

/* FinishTopology(false,true) -> ComputeBounds VG.cpp:1792-1816: the first
 * leaf seeds the box, then every active leaf widens it */
static void compute_bounds(or_volume* v) {
    int64_t rg0 = v->range[0];
    or_node* c0 = &v->nodes[0][0];
    float vmin[3] = {(float)c0->mPos[0], (float)c0->mPos[1], (float)c0->mPos[2]};
    float vmax[3] = {vmin[0], vmin[1], vmin[2]};
    for (int64_t i = 0; i < v->ncnt[0]; i++) {
        or_node* c = &v->nodes[0][i];
        if (!c->mFlags) continue;
        for (int a = 0; a < 3; a++) {
            if (c->mPos[a] < vmin[a]) vmin[a] = (float)c->mPos[a];
            if (c->mPos[a] + rg0 > vmax[a]) vmax[a] = (float)(c->mPos[a] + rg0);
        }
    }
    memcpy(v->obj_min, vmin, sizeof(vmin));
    memcpy(v->obj_max, vmax, sizeof(vmax));
}

int or_rebuild_topology(or_volume* v, const float* pos, int64_t n, int64_t stride, float radius,
                        const float orig[3]) {
    if (!v || !v->configured || (!pos && n > 0) || n < 0) return OR_ERR_ARG;
    if (stride == 0) stride = 12;
    /* GetBoundingBox VG.cpp:5404-5467 over valid points (D1) */
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t nvalid = 0;
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), orig, w)) continue;
        nvalid++;
        for (int a = 0; a < 3; a++) {
            if (w[a] < mn[a]) mn[a] = w[a];
            if (w[a] > mx[a]) mx[a] = w[a];
        }
    }
    if (nvalid == 0) return OR_ERR_NOPOINTS;
    float pmin[3], pmax[3];
    for (int a = 0; a < 3; a++) {
        pmin[a] = mn[a] - 1;
        pmax[a] = mx[a] + 2;
        pmin[a] = 0 > pmin[a] ? 0 : pmin[a];
        pmax[a] = 0 > pmax[a] ? 0 : pmax[a];
    }
    /* DetermineDepth VG.cpp:821-837 with GetCoveringNode VG.cpp:2915-2931 */
    int ipmin[3] = {(int)pmin[0], (int)pmin[1], (int)pmin[2]};
    int ipmax[3] = {(int)pmax[0], (int)pmax[1], (int)pmax[2]};
    int depth = -1, rootp[3] = {0, 0, 0};
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
            depth = lev;
            rootp[0] = (int)a0[0]; rootp[1] = (int)a0[1]; rootp[2] = (int)a0[2];
            break;
        }
    }
    if (depth <= 0 || depth >= v->levs) return OR_ERR_DEPTH;  /* VG.cpp:1187-1194 */
    /* u16 key range (PT.cuh:377-380) */
    for (int a = 0; a < 3; a++)
        if ((int64_t)ipmax[a] / v->range[0] > 65535) return OR_ERR_RANGE;

    memcpy(v->pos_min, pmin, sizeof(pmin));
    memcpy(v->pos_max, pmax, sizeof(pmax));
    /* Clear VG.cpp:2487-2503 */
    for (int l = 0; l < OR_MAXLEV; l++) { v->ncnt[l] = 0; v->ccnt[l] = 0; }
    free_sc(v);
    v->depth = depth;
    memcpy(v->root_pos, rootp, sizeof(rootp));

    /* ---- phase 1: ActivateBricksGPU VG.cpp:1229-1413 ---- */
    uint64_t* keys = (uint64_t*)malloc((size_t)(nvalid * depth + 1) * sizeof(uint64_t));
    int64_t nk = 0;
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), orig, w)) continue;
        for (int lev = 0; lev < depth; lev++) {   /* gvdbCalcBrickId PT.cuh:352-382 */
            int rg = (int)v->range[lev];
            int bx = cover_axis(w[0], rg), by = cover_axis(w[1], rg), bz = cover_axis(w[2], rg);
            keys[nk++] = make_key(lev, bx / rg, by / rg, bz / rg);
        }
    }
    qsort(keys, (size_t)nk, sizeof(uint64_t), cmp_u64);
    int64_t nu = unique_u64(keys, nk);

    int64_t root = pool_alloc_node(v, depth);
    int64_t rcl = pool_alloc_clist(v, depth);
    setup_node(v, depth, root, rootp[0], rootp[1], rootp[2]);
    v->nodes[depth][root].mFlags = 1;
    v->nodes[depth][root].mChildList = elem(1, depth, rcl);
    alloc_from_keys(v, keys, nu);
    compute_top_lev(v);
    for (int lev = depth - 1; lev >= 0; lev--) link_level(v, lev);

    /* ---- phase 2: ActivateExtraBricksGPU VG.cpp:931-1016 ---- */
    if (radius >= 1) {
        int64_t cap = 1024, ne = 0;
        uint64_t* ek = (uint64_t*)malloc((size_t)cap * sizeof(uint64_t));
        int64_t rtop = v->range[depth];
        for (int64_t i = 0; i < n; i++) {
            float w[3];
            if (!point_valid(pt(pos, i, stride), orig, w)) continue;
            /* gvdbCalcExtraBrickId PT.cuh:467-512 */
            int bmin[3], bmax[3];
            for (int a = 0; a < 3; a++) {
                bmin[a] = (int)((float)(int)w[a] - radius);
                bmax[a] = (int)((float)(int)w[a] + 2.0f * radius);
            }
            for (int lev = 0; lev < depth; lev++) {
                int rres = (int)v->range[lev];
                int nd[3];
                for (nd[0] = bmin[0] / rres * rres; nd[0] <= bmax[0] / rres * rres; nd[0] += rres)
                    for (nd[1] = bmin[1] / rres * rres; nd[1] <= bmax[1] / rres * rres; nd[1] += rres)
                        for (nd[2] = bmin[2] / rres * rres; nd[2] <= bmax[2] / rres * rres; nd[2] += rres) {
                            int inside = 1;   /* D3: clip to the root node */
                            for (int a = 0; a < 3; a++)
                                if (nd[a] < rootp[a] || (int64_t)nd[a] >= rootp[a] + rtop) inside = 0;
                            if (!inside) continue;
                            uint64_t k = make_key(lev, nd[0] / rres, nd[1] / rres, nd[2] / rres);
                            if (key_in(keys, nu, k)) continue;   /* D3: getPosParent != UNDEF */
                            if (ne >= cap) {
                                cap *= 2;
                                ek = (uint64_t*)realloc(ek, (size_t)cap * sizeof(uint64_t));
                            }
                            ek[ne++] = k;
                        }
            }
        }
        qsort(ek, (size_t)ne, sizeof(uint64_t), cmp_u64);
        int64_t neu = unique_u64(ek, ne);
        alloc_from_keys(v, ek, neu);
        compute_top_lev(v);
        for (int lev = depth - 1; lev >= 0; lev--) link_level(v, lev);
        free(ek);
    }
    free(keys);

    compute_bounds(v);
    v->atlas_used = 0;
    return OR_OK;
}

/* getPosParent ND.cuh:303-348 (no bitmasks): the node of level stop_lev
 * holding pos, found by descending the child lists from the top node;
 * -1 when a list on the way has no entry */
static int64_t pos_parent(const or_volume* v, const int pos[3], int stop_lev) {
    if (v->ncnt[v->top_lev] < 1) return -1;
    const or_node* nd = &v->nodes[v->top_lev][0];
    int l = nd->mLev;
    if (l == 0) return 0;
    while (l > 0) {
        int res = v->res[l];
        int64_t pin[3];
        pos_in_node(pos, nd->mPos, res, v->range[l], pin);
        int64_t bit = (pin[2] * res + pin[1]) * res + pin[0];
        int64_t c = get_child(v, nd, bit);
        l--;
        if (c < 0) return -1;
        if (l == stop_lev) return c;
        if (c >= v->ncnt[l]) return -1;
        nd = &v->nodes[l][c];
    }
    return -1;
}

/* the parent's child-list slot of a node at level lev, or -1 (link/delink
 * position arithmetic ND.cuh:414-428) */
static int64_t parent_slot(const or_volume* v, int lev, const or_node* node, int64_t p) {
    const or_node* pn = &v->nodes[lev + 1][p];
    int res = v->res[lev + 1];
    int pos[3] = {node->mPos[0], node->mPos[1], node->mPos[2]};
    int64_t pin[3];
    pos_in_node(pos, pn->mPos, res, v->range[lev + 1], pin);
    for (int a = 0; a < 3; a++)
        if (pin[a] < 0 || pin[a] >= res) return -1;
    int64_t len = ipow3(res);
    int64_t flat = (int64_t)(pn->mChildList >> 16) * len + (pin[2] * res + pin[1]) * res + pin[0];
    if (flat < 0 || flat >= v->ccnt[lev + 1] * len) return -1;   /* D2 */
    return flat;
}

/* ActivateIncreBricksGPU VG.cpp:1020-1155.  accum = 1 is AccumulateTopology
 * (VG.cpp:1158-1178): existing leaves keep their flags; accum = 0 is the
 * incremental RebuildTopology (VG.cpp:1215-1216): leaves no point reaches
 * are deactivated and delinked.  Every valid point's box
 * [int(w) - r, int(w) + r] (gvdbCalcIncreExtraBrickId PT.cuh:420-465) is
 * walked at every level below the root: reachable nodes are marked (leaves),
 * missing ones become keys; the sorted unique keys are appended to the
 * pools; the flagged nodes are linked top-down (ND.cuh:470-505), then the
 * unflagged leaves and the childless inner nodes are delinked bottom-up
 * (gvdbDelinkLeafBricks / gvdbDelinkBricks ND.cuh:401-467).  With no
 * missing node nothing changes (VG.cpp:1066).  Deviations: the tree's own
 * depth and root are kept, so every valid point must lie inside the root
 * (else OR_ERR_RANGE and no change); box cells outside the root are not
 * emitted (D3). */
int or_incremental_topology(or_volume* v, const float* pos, int64_t n, int64_t stride, float radius,
                            const float orig[3], int accum) {
    if (!v || !v->configured || (!pos && n > 0) || n < 0 || !(radius >= 0.0f) || !isfinite(radius)) return OR_ERR_ARG;
    if (v->ncnt[0] == 0 || v->depth <= 0) return OR_ERR_STATE;   /* VG.cpp:1022 */
    if (v->top_lev != v->depth) return OR_ERR_STATE;
    if (stride == 0) stride = 12;
    const int depth = v->depth;
    const int64_t rtop = v->range[depth];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t nvalid = 0;
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), orig, w)) continue;
        nvalid++;
        for (int a = 0; a < 3; a++) {
            if (w[a] < mn[a]) mn[a] = w[a];
            if (w[a] > mx[a]) mx[a] = w[a];
        }
    }
    if (nvalid == 0) return OR_ERR_NOPOINTS;
    for (int a = 0; a < 3; a++)
        if (mn[a] < (float)v->root_pos[a] || mx[a] >= (float)((int64_t)v->root_pos[a] + rtop)) return OR_ERR_RANGE;
    /* GetBoundingBox VG.cpp:5404-5467 */
    for (int a = 0; a < 3; a++) {
        float lo = mn[a] - 1, hi = mx[a] + 2;
        v->pos_min[a] = 0 > lo ? 0 : lo;
        v->pos_max[a] = 0 > hi ? 0 : hi;
    }
    /* node markers (VG.cpp:1026-1041) */
    int64_t nleaf = v->ncnt[0];
    uint8_t* marker = (uint8_t*)malloc((size_t)nleaf);
    for (int64_t k = 0; k < nleaf; k++) marker[k] = accum ? v->nodes[0][k].mFlags : 0;
    int64_t cap = 1024, ne = 0;
    uint64_t* keys = (uint64_t*)malloc((size_t)cap * sizeof(uint64_t));
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), orig, w)) continue;
        int bmin[3], bmax[3];
        for (int a = 0; a < 3; a++) {
            bmin[a] = (int)((float)(int)w[a] - radius);
            bmax[a] = (int)((float)(int)w[a] + radius);
        }
        for (int lev = 0; lev < depth; lev++) {
            int rres = (int)v->range[lev];
            int nd[3];
            for (nd[0] = bmin[0] / rres * rres; nd[0] <= bmax[0] / rres * rres; nd[0] += rres)
                for (nd[1] = bmin[1] / rres * rres; nd[1] <= bmax[1] / rres * rres; nd[1] += rres)
                    for (nd[2] = bmin[2] / rres * rres; nd[2] <= bmax[2] / rres * rres; nd[2] += rres) {
                        int inside = 1;   /* D3 */
                        for (int a = 0; a < 3; a++)
                            if (nd[a] < v->root_pos[a] || (int64_t)nd[a] >= v->root_pos[a] + rtop) inside = 0;
                        if (!inside) continue;
                        int64_t id = pos_parent(v, nd, lev);
                        if (id < 0) {
                            if (ne >= cap) {
                                cap *= 2;
                                keys = (uint64_t*)realloc(keys, (size_t)cap * sizeof(uint64_t));
                            }
                            keys[ne++] = make_key(lev, nd[0] / rres, nd[1] / rres, nd[2] / rres);
                        } else if (lev == 0 && id < nleaf) {
                            marker[id] = 1;
                        }
                    }
        }
    }
    if (ne == 0) {   /* no new node: nothing changes (VG.cpp:1066) */
        free(keys);
        free(marker);
        return OR_OK;
    }
    qsort(keys, (size_t)ne, sizeof(uint64_t), cmp_u64);
    int64_t nu = unique_u64(keys, ne);
    for (int64_t k = 0; k < nleaf; k++) v->nodes[0][k].mFlags = marker[k];   /* VG.cpp:1078-1085 */
    free(marker);
    alloc_from_keys(v, keys, nu);   /* VG.cpp:1087-1111 */
    free(keys);
    compute_top_lev(v);             /* PrepareVDBPartially */
    for (int lev = depth - 1; lev >= 0; lev--) link_level(v, lev);
    /* gvdbDelinkLeafBricks ND.cuh:401-431 */
    for (int64_t k = 0; k < v->ncnt[0]; k++) {
        or_node* node = &v->nodes[0][k];
        if (node->mFlags || node->mParent == OR_ID_UNDEF64) continue;
        int p3[3] = {node->mPos[0], node->mPos[1], node->mPos[2]};
        int64_t p = get_parent(v, 1, p3);
        if (p < 0 || p >= v->ncnt[1]) continue;
        int64_t slot = parent_slot(v, 0, node, p);
        node->mParent = OR_ID_UNDEF64;
        if (slot >= 0) v->clist[1][slot] = OR_ID_UNDEF64;
    }
    /* gvdbDelinkBricks ND.cuh:433-467, one level after the other */
    for (int lev = 1; lev < depth; lev++) {
        int64_t len = ipow3(v->res[lev]);
        for (int64_t k = 0; k < v->ncnt[lev]; k++) {
            or_node* node = &v->nodes[lev][k];
            if (node->mParent == OR_ID_UNDEF64) { node->mFlags = 0; continue; }
            const uint64_t* own = v->clist[lev] + k * len;
            int any = 0;
            for (int64_t c = 0; c < len && !any; c++) any = own[c] != OR_ID_UNDEF64;
            if (any) { node->mFlags = 1; continue; }
            node->mFlags = 0;
            int p3[3] = {node->mPos[0], node->mPos[1], node->mPos[2]};
            int64_t p = get_parent(v, lev + 1, p3);
            if (p < 0 || p >= v->ncnt[lev + 1]) continue;
            int64_t slot = parent_slot(v, lev, node, p);
            node->mParent = OR_ID_UNDEF64;
            if (slot >= 0) v->clist[lev + 1][slot] = OR_ID_UNDEF64;
        }
    }
    free_sc(v);
    compute_bounds(v);
    v->atlas_used = 0;
    return OR_OK;
}

int or_update_atlas(or_volume* v) {
    if (!v || !v->configured) return OR_ERR_ARG;
    int any = 0;
    for (int c = 0; c < OR_MAXCHAN; c++) any |= v->has_chan[c];
    if (!any) return OR_ERR_STATE;
    /* the pool's used count: every leaf after a rebuild, the active ones
     * after an incremental update (UpdateUsed VG.cpp:1145-1149) */
    int64_t used = 0;
    for (int64_t k = 0; k < v->ncnt[0]; k++) used += v->nodes[0][k].mFlags ? 1 : 0;
    int64_t amax = (int64_t)v->subdim[0] * v->subdim[1] * v->subdim[2];
    if (used > amax) {   /* VG.cpp:2645-2651, AL.cpp:622-650 */
        int sz = (int)ceil(used / (float)(v->subdim[0] * v->subdim[1]));
        if (sz > v->subdim[2]) atlas_resize(v, sz);
    }
    int br = brick_res(v);
    int sx = v->subdim[0], sy = v->subdim[1];
    int64_t id = 0;
    for (int64_t k = 0; k < v->ncnt[0]; k++) {   /* AtlasAlloc/getAtlasPos AL.cpp:692-715 */
        or_node* nd = &v->nodes[0][k];
        if (!nd->mFlags) continue;
        int64_t a2 = (int64_t)sx * sy;
        int64_t z = id / a2, rem = id - z * a2;
        int64_t y = rem / sx, x = rem - y * sx;
        nd->mValue[0] = (int)(x * br + v->apron);
        nd->mValue[1] = (int)(y * br + v->apron);
        nd->mValue[2] = (int)(z * br + v->apron);
        id++;
    }
    v->atlas_used = id;
    /* AllocateAtlasMap + ClearMapping + AssignMapping VG.cpp:2559-2588,2676-2689 */
    int64_t mc = (int64_t)sx * sy * v->subdim[2];
    free(v->amap);
    v->amap = (or_atlas_node*)malloc((size_t)mc * sizeof(or_atlas_node));
    v->amap_cnt = mc;
    for (int64_t i = 0; i < mc; i++) {
        v->amap[i].mPos[0] = v->amap[i].mPos[1] = v->amap[i].mPos[2] = (int32_t)OR_ID_UNDEFL;
        v->amap[i].mLeafID = (int32_t)OR_ID_UNDEFL;
    }
    for (int64_t k = 0; k < v->ncnt[0]; k++) {
        or_node* nd = &v->nodes[0][k];
        if (!nd->mFlags) continue;
        int64_t ix = nd->mValue[0] / br, iy = nd->mValue[1] / br, iz = nd->mValue[2] / br;
        int64_t slot = (iz * sy + iy) * sx + ix;
        memcpy(v->amap[slot].mPos, nd->mPos, sizeof(nd->mPos));
        v->amap[slot].mLeafID = (int32_t)k;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* tree lookups                                                        */
/* ------------------------------------------------------------------ */

/* getPosLeafParent ND.cuh:255-288 + D4 containment */
static int64_t pos_leaf(const or_volume* v, const int pos[3]) {
    if (v->ncnt[v->top_lev] < 1) return -1;
    const or_node* nd = &v->nodes[v->top_lev][0];
    int l = nd->mLev;
    int64_t id = -1;
    if (l == 0) {
        id = 0;
    } else {
        while (1) {
            int res = v->res[l];
            int64_t pin[3];
            pos_in_node(pos, nd->mPos, res, v->range[l], pin);
            if (pin[0] > res || pin[1] > res || pin[2] > res) return -1;
            int64_t bit = (pin[2] * res + pin[1]) * res + pin[0];
            int64_t c = get_child(v, nd, bit);
            if (c < 0) return -1;
            l--;
            if (c >= v->ncnt[l]) return -1;
            if (l == 0) { id = c; break; }
            nd = &v->nodes[l][c];
        }
    }
    const or_node* lf = &v->nodes[0][id];
    for (int a = 0; a < 3; a++)
        if (pos[a] < lf->mPos[a] || (int64_t)pos[a] >= lf->mPos[a] + v->res[0]) return -1;
    return id;
}

/* getNodeAtPoint ND.cuh:199-226,243-253 (float descent) + D4 */
static int64_t point_leaf(const or_volume* v, const float pos[3]) {
    int lev = v->top_lev;
    if (v->ncnt[lev] < 1) return -1;
    const or_node* node = &v->nodes[lev][0];
    int64_t id = 0;
    float vmin[3] = {(float)node->mPos[0], (float)node->mPos[1], (float)node->mPos[2]};
    while (lev > 0) {
        float rg = (float)v->range[lev];
        if (pos[0] < vmin[0] || pos[1] < vmin[1] || pos[2] < vmin[2] ||
            pos[0] >= vmin[0] + rg || pos[1] >= vmin[1] + rg || pos[2] >= vmin[2] + rg)
            return -1;
        float vdel = (float)v->range[lev] / (float)v->res[lev];
        int p0 = (int)((pos[0] - vmin[0]) / vdel), p1 = (int)((pos[1] - vmin[1]) / vdel),
            p2 = (int)((pos[2] - vmin[2]) / vdel);
        int64_t b = (((int64_t)p2 << v->logdim[lev]) + p1) << v->logdim[lev];
        b += p0;
        lev--;
        int64_t c = get_child(v, node, b);
        if (c < 0 || c >= v->ncnt[lev]) return -1;
        id = c;
        node = &v->nodes[lev][c];
        vmin[0] = (float)node->mPos[0]; vmin[1] = (float)node->mPos[1]; vmin[2] = (float)node->mPos[2];
    }
    float r0 = (float)v->res[0];
    for (int a = 0; a < 3; a++)
        if (!(pos[a] >= vmin[a] && pos[a] < vmin[a] + r0)) return -1;
    return id;
}

static int64_t atlas_index(const or_volume* v, int64_t x, int64_t y, int64_t z) {
    int br = brick_res(v);
    int64_t ax = (int64_t)v->subdim[0] * br, ay = (int64_t)v->subdim[1] * br;
    return (z * ay + y) * ax + x;
}

/* ------------------------------------------------------------------ */
/* points -> subcells -> gather                                        */
/* ------------------------------------------------------------------ */

/* FP16 list entry of gvdbInsertSubcell_fp16 (PT.cuh:206-209) read back as
 * gvdbGatherLevelSet_fp16 does (PT.cuh:794-796): ushort((w - mn) / rg *
 * 65535) (truncated, clamped to [0, 65535], NaN -> 0), then q / 65535 * rg +
 * mn; FP32 throughout, in that order. */
static float fp16_round_trip(float w, float mn, float rg) {
    float t = ((w - mn) / rg) * 65535.0f;
    uint32_t q = !(t > 0.0f) ? 0u : (t >= 65535.0f ? 65535u : (uint32_t)t);
    return ((float)q / 65535.0f) * rg + mn;
}

static int insert_sc_impl(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride, float radius,
                          const float trans[3], int64_t* out_len, int fp16);

int or_insert_points_subcell(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride,
                             float radius, const float trans[3], int64_t* out_len) {
    return insert_sc_impl(v, sc, pos, n, stride, radius, trans, out_len, 0);
}

/* InsertPointsSubcell_FP16 VG.cpp:5094-5189: the same lists, positions kept
 * as 16-bit fractions of GetBoundingBox's box (mPosMin, mPosRange =
 * mPosMax - mPosMin, VG.cpp:5452-5462); or_gather then reads them back, which
 * with mode 1 and no accumulate is GatherLevelSet_FP16 (PT.cuh:775-811). */
int or_insert_points_subcell_fp16(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride,
                                  float radius, const float trans[3], int64_t* out_len) {
    return insert_sc_impl(v, sc, pos, n, stride, radius, trans, out_len, 1);
}

static int insert_sc_impl(or_volume* v, int sc, const float* pos, int64_t n, int64_t stride, float radius,
                          const float trans[3], int64_t* out_len, int fp16) {
    if (!v || !v->configured || sc <= 0 || (!pos && n > 0)) return OR_ERR_ARG;
    if (stride == 0) stride = 12;
    int64_t nb = v->ncnt[0];
    if (nb == 0) return OR_ERR_STATE;   /* VG.cpp:4973-4977 */
    if (v->res[0] % sc) return OR_ERR_ARG;
    free_sc(v);
    int scdim = v->res[0] / sc;
    int per = scdim * scdim * scdim;
    v->sc_size = sc; v->sc_dim = scdim; v->sc_per_brick = per;
    /* gvdbSetFlagSubcell PT.cuh:109-118 + PrefixSum -> sc_mapping */
    int32_t* mapping = (int32_t*)malloc((size_t)nb * sizeof(int32_t));
    int32_t acc = 0;
    for (int64_t i = 0; i < nb; i++) {
        mapping[i] = acc;
        acc += (v->nodes[0][i].mParent == OR_ID_UNDEF64) ? 0 : 1;
    }
    int64_t nsc = (int64_t)per * nb;
    v->num_sc = nsc;
    v->sc_cnt = (int32_t*)calloc((size_t)nsc, sizeof(int32_t));
    v->sc_off = (int64_t*)calloc((size_t)nsc + 1, sizeof(int64_t));
    /* two passes: gvdbCountSubcell PT.cuh:275-315 then gvdbInsertSubcell PT.cuh:218-273 */
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            int64_t run = 0;
            for (int64_t s = 0; s < nsc; s++) { v->sc_off[s] = run; run += v->sc_cnt[s]; }
            v->sc_off[nsc] = run;
            v->sc_len = run;
            v->sc_pnt = (float*)malloc((size_t)(run * 3 + 3) * sizeof(float));
            /* point indices only feed the colour mean: kept while a uchar4 channel exists */
            int keep = 0;
            for (int c = 0; c < OR_MAXCHAN; c++) keep |= v->has_chan[c] && v->chan_type[c] == OR_T_UCHAR4;
            v->sc_pidx = keep ? (int64_t*)malloc((size_t)(run + 1) * sizeof(int64_t)) : NULL;
            memset(v->sc_cnt, 0, (size_t)nsc * sizeof(int32_t));
        }
        for (int64_t i = 0; i < n; i++) {
            float w[3];
            if (!point_valid(pt(pos, i, stride), trans, w)) continue;
            int lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                lo[a] = ((int)(w[a] - radius) / sc) * sc;
                hi[a] = ((int)(w[a] + 1 + radius) / sc) * sc;
            }
            int s3[3];
            for (s3[0] = lo[0]; s3[0] <= hi[0]; s3[0] += sc)
                for (s3[1] = lo[1]; s3[1] <= hi[1]; s3[1] += sc)
                    for (s3[2] = lo[2]; s3[2] <= hi[2]; s3[2] += sc) {
                        int64_t leaf = pos_leaf(v, s3);
                        if (leaf < 0) continue;
                        const or_node* lf = &v->nodes[0][leaf];
                        int px = (s3[0] - lf->mPos[0]) / sc, py = (s3[1] - lf->mPos[1]) / sc,
                            pz = (s3[2] - lf->mPos[2]) / sc;
                        int loc = (pz * scdim + py) * scdim + px;
                        int64_t sidx = (int64_t)mapping[leaf] * per + loc;
                        if (sidx < 0 || sidx >= nsc) continue;
                        if (pass == 1) {
                            int64_t e = v->sc_off[sidx] + v->sc_cnt[sidx];
                            for (int a = 0; a < 3; a++)
                                v->sc_pnt[e * 3 + a] =
                                    fp16 ? fp16_round_trip(w[a], v->pos_min[a], v->pos_max[a] - v->pos_min[a]) : w[a];
                            if (v->sc_pidx) v->sc_pidx[e] = i;
                        }
                        v->sc_cnt[sidx]++;
                    }
        }
    }
    /* gvdbCalcSubcellPos PT.cuh:677-694 */
    v->sc_nid = (int32_t*)malloc((size_t)nsc * sizeof(int32_t));
    v->sc_pos = (int32_t*)malloc((size_t)nsc * 3 * sizeof(int32_t));
    for (int64_t s = 0; s < nsc; s++) v->sc_nid[s] = -1;
    for (int64_t i = 0; i < nb; i++) {
        const or_node* nd = &v->nodes[0][i];
        if (nd->mParent == OR_ID_UNDEF64) continue;
        for (int s = 0; s < per; s++) {
            int64_t idx = (int64_t)mapping[i] * per + s;
            v->sc_pos[idx * 3 + 0] = nd->mPos[0] + s % scdim * sc;
            v->sc_pos[idx * 3 + 1] = nd->mPos[1] + (s / scdim) % scdim * sc;
            v->sc_pos[idx * 3 + 2] = nd->mPos[2] + (s / scdim / scdim) % scdim * sc;
            v->sc_nid[idx] = (int32_t)i;
        }
    }
    free(mapping);
    if (out_len) *out_len = v->sc_len;
    return OR_OK;
}

static float dist2f(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

int or_gather(or_volume* v, int mode, float radius, int chan, int accum) {
    if (!v || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan]) return OR_ERR_ARG;
    if (!v->sc_cnt || !v->sc_nid) return OR_ERR_STATE;
    float* A = v->atlas[chan];
    int sc = v->sc_size;
    for (int64_t s = 0; s < v->num_sc; s++) {
        int32_t nid = v->sc_nid[s];
        if (nid < 0) continue;
        const or_node* nd = &v->nodes[0][nid];
        for (int tz = 0; tz < sc; tz++)
            for (int ty = 0; ty < sc; ty++)
                for (int tx = 0; tx < sc; tx++) {
                    int wp[3] = {v->sc_pos[s * 3] + tx, v->sc_pos[s * 3 + 1] + ty, v->sc_pos[s * 3 + 2] + tz};
                    int64_t vx = nd->mValue[0] + (wp[0] - nd->mPos[0]);
                    int64_t vy = nd->mValue[1] + (wp[1] - nd->mPos[1]);
                    int64_t vz = nd->mValue[2] + (wp[2] - nd->mPos[2]);
                    int64_t ai = atlas_index(v, vx, vy, vz);
                    float val;
                    if (mode == 1) val = accum ? A[ai] : 3.0f;   /* PT.cuh:753-756 */
                    else val = accum ? A[ai] : 0.0f;              /* PT.cuh:714-717 */
                    for (int64_t j = v->sc_off[s]; j < v->sc_off[s] + v->sc_cnt[s]; j++) {
                        float jx = v->sc_pnt[j * 3] - (float)wp[0];
                        float jy = v->sc_pnt[j * 3 + 1] - (float)wp[1];
                        float jz = v->sc_pnt[j * 3 + 2] - (float)wp[2];
                        float c = sqrtf(dist2f(jx, jy, jz));
                        if (mode == 1) val = fminf(val, c - radius);
                        else val = fmaxf(val, radius - c);
                    }
                    A[ai] = val;
                }
    }
    return OR_OK;
}

int or_gather_f64(or_volume* v, int mode, float radius, int accum, double* out) {
    if (!v || !out) return OR_ERR_ARG;
    if (!v->sc_cnt || !v->sc_nid) return OR_ERR_STATE;
    if (accum) return OR_ERR_ARG;
    int sc = v->sc_size;
    int r0 = v->res[0];
    for (int64_t s = 0; s < v->num_sc; s++) {
        int32_t nid = v->sc_nid[s];
        if (nid < 0) continue;
        const or_node* nd = &v->nodes[0][nid];
        for (int tz = 0; tz < sc; tz++)
            for (int ty = 0; ty < sc; ty++)
                for (int tx = 0; tx < sc; tx++) {
                    int wp[3] = {v->sc_pos[s * 3] + tx, v->sc_pos[s * 3 + 1] + ty, v->sc_pos[s * 3 + 2] + tz};
                    double val = mode == 1 ? 3.0 : 0.0;
                    for (int64_t j = v->sc_off[s]; j < v->sc_off[s] + v->sc_cnt[s]; j++) {
                        double jx = (double)v->sc_pnt[j * 3] - wp[0];
                        double jy = (double)v->sc_pnt[j * 3 + 1] - wp[1];
                        double jz = (double)v->sc_pnt[j * 3 + 2] - wp[2];
                        double c = sqrt(jx * jx + jy * jy + jz * jz);
                        if (mode == 1) val = fmin(val, c - (double)radius);
                        else val = fmax(val, (double)radius - c);
                    }
                    int lx = wp[0] - nd->mPos[0], ly = wp[1] - nd->mPos[1], lz = wp[2] - nd->mPos[2];
                    out[(int64_t)nid * r0 * r0 * r0 + ((int64_t)lz * r0 + ly) * r0 + lx] = val;
                }
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* apron                                                               */
/* ------------------------------------------------------------------ */

int or_update_apron(or_volume* v, int chan, float boundval) {
    if (!v || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan]) return OR_ERR_ARG;
    if (v->apron == 0) return OR_OK;        /* VG.cpp:4420 */
    if (v->apron != 1) return OR_ERR_ARG;   /* OP.cuh kernels assume apron 1 (VG.cpp:4450) */
    if (!v->amap) return OR_ERR_STATE;
    uint32_t* A = (uint32_t*)v->atlas[chan];   /* the apron copies 4-byte voxels of either type */
    const uint32_t bound = bound_word(v->chan_type[chan], boundval);
    int br = brick_res(v);
    int sx = v->subdim[0], sy = v->subdim[1];
    /* UpdateApron<T> OP.cuh:73-126 runs over the atlas bricks: every brick
     * the atlas map assigns to a leaf (deactivated leaves own none) */
    for (int64_t slot = 0; slot < v->amap_cnt; slot++) {
        const int32_t b = v->amap[slot].mLeafID;
        if (b == (int32_t)OR_ID_UNDEFL || b < 0 || b >= v->ncnt[0]) continue;
        const or_node* node = &v->nodes[0][b];
        for (int side = 0; side < 6; side++)
            for (int su = 0; su < br; su++)
                for (int sv = 0; sv < br; sv++) {
                    int bv[3];
                    switch (side) {
                        case 0: bv[0] = 0; bv[1] = su; bv[2] = sv; break;
                        case 1: bv[0] = su; bv[1] = 0; bv[2] = sv; break;
                        case 2: bv[0] = su; bv[1] = sv; bv[2] = 0; break;
                        case 3: bv[0] = br - 1; bv[1] = su; bv[2] = sv; break;
                        case 4: bv[0] = su; bv[1] = br - 1; bv[2] = sv; break;
                        default: bv[0] = su; bv[1] = sv; bv[2] = br - 1; break;
                    }
                    int64_t av[3];
                    for (int a = 0; a < 3; a++) av[a] = (int64_t)bv[a] + node->mValue[a] - 1;
                    if (av[0] < 0 || av[1] < 0 || av[2] < 0) continue;
                    /* getAtlasToWorld ND.cuh:145-153 */
                    int64_t bx = av[0] / br, by = av[1] / br, bz = av[2] / br;
                    int64_t bid = (bz * sy + by) * sx + bx;
                    if (bid >= v->amap_cnt || v->amap[bid].mLeafID == (int32_t)OR_ID_UNDEFL) continue;
                    float wpos[3];
                    for (int a = 0; a < 3; a++)
                        wpos[a] = (float)v->amap[bid].mPos[a] + (float)(int)(av[a] % br - v->apron) + 0.5f;
                    int64_t leaf = point_leaf(v, wpos);
                    uint32_t value;
                    if (leaf < 0) {
                        value = bound;
                    } else {
                        const or_node* ln = &v->nodes[0][leaf];
                        float ox = (float)ln->mValue[0] + (wpos[0] - (float)ln->mPos[0]);
                        float oy = (float)ln->mValue[1] + (wpos[1] - (float)ln->mPos[1]);
                        float oz = (float)ln->mValue[2] + (wpos[2] - (float)ln->mPos[2]);
                        value = A[atlas_index(v, (int64_t)(uint32_t)ox, (int64_t)(uint32_t)oy, (int64_t)(uint32_t)oz)];
                    }
                    A[atlas_index(v, av[0], av[1], av[2])] = value;
                }
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* scatter path                                                        */
/* ------------------------------------------------------------------ */

int or_insert_points(or_volume* v, const float* pos, int64_t n, int64_t stride, const float trans[3],
                     int32_t* pnode, int32_t* poff) {
    if (!v || (!pos && n > 0) || !pnode || !poff) return OR_ERR_ARG;
    if (stride == 0) stride = 12;
    int64_t nb = v->ncnt[0];
    int32_t* gcnt = (int32_t*)calloc((size_t)(nb + 1), sizeof(int32_t));
    for (int64_t i = 0; i < n; i++) {   /* gvdbInsertPoints PT.cuh:19-36 */
        float w[3];
        pnode[i] = (int32_t)OR_ID_UNDEFL;
        poff[i] = 0;
        if (!point_valid(pt(pos, i, stride), trans, w)) continue;
        int64_t leaf = point_leaf(v, w);
        if (leaf < 0) continue;
        pnode[i] = (int32_t)leaf;
        poff[i] = gcnt[leaf]++;
    }
    free(gcnt);
    return OR_OK;
}

/* distFunc PT.cuh:516-523 */
static float dist_func(const float a[3], float bx, float by, float bz, float r) {
    bx -= a[0]; by -= a[1]; bz -= a[2];
    float c = dist2f(bx, by, bz) / (r * r);
    float t = 3.0f - c;
    t = c * t;
    t = -3.0f + t;
    t = c * t;
    return 1.0f + t;
}

int or_scatter_density(or_volume* v, const float* pos, int64_t n, int64_t stride, float radius, float amp,
                       const float trans[3], int expand, int chan) {
    (void)amp;   /* unused by the reference kernel (PT.cuh:525-573) */
    if (!v || (!pos && n > 0) || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan]) return OR_ERR_ARG;
    if (stride == 0) stride = 12;
    float* A = v->atlas[chan];
    int r0 = v->res[0];
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), trans, w)) continue;
        int64_t leaf = point_leaf(v, w);   /* pnode from InsertPoints */
        if (leaf < 0) continue;
        const or_node* nd = &v->nodes[0][leaf];
        float p[3] = {w[0] - (float)nd->mPos[0], w[1] - (float)nd->mPos[1], w[2] - (float)nd->mPos[2]};
        float pi[3] = {(float)(int)p[0], (float)(int)p[1], (float)(int)p[2]};
        if (pi[0] < 0 || pi[1] < 0 || pi[2] < 0 || pi[0] >= r0 || pi[1] >= r0 || pi[2] >= r0) continue;
        int64_t q[3] = {(int64_t)pi[0] + nd->mValue[0], (int64_t)pi[1] + nd->mValue[1], (int64_t)pi[2] + nd->mValue[2]};
        static const int off[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
        int cnt = expand ? 7 : 1;
        for (int k = 0; k < cnt; k++) {
            int64_t ai = atlas_index(v, q[0] + off[k][0], q[1] + off[k][1], q[2] + off[k][2]);
            A[ai] = A[ai] + dist_func(p, pi[0] + off[k][0], pi[1] + off[k][1], pi[2] + off[k][2], radius);
        }
    }
    return OR_OK;
}

/* Leaf lookup for an integer voxel coordinate. */
static int64_t voxel_leaf(const or_volume* v, int64_t x, int64_t y, int64_t z) {
    float c[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
    return point_leaf(v, c);
}

/* Weighted cubic splat, extension mode (iii): rho(v) = sum_p w_p max(0,1-|v-p|^2/r^2)^3
 * over integer voxel coordinates v of active leaves, summed in ascending point order. */
int or_splat_cubic(or_volume* v, const float* pos, int64_t n, int64_t stride, const float* weights,
                   float radius, const float trans[3], int chan, int accum) {
    if (!v || (!pos && n > 0) || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan]) return OR_ERR_ARG;
    if (!v->amap) return OR_ERR_STATE;
    if (stride == 0) stride = 12;
    float* A = v->atlas[chan];
    int r0 = v->res[0];
    if (!accum) {
        for (int64_t b = 0; b < v->ncnt[0]; b++) {
            const or_node* nd = &v->nodes[0][b];
            for (int z = 0; z < r0; z++)
                for (int y = 0; y < r0; y++)
                    for (int x = 0; x < r0; x++)
                        A[atlas_index(v, nd->mValue[0] + x, nd->mValue[1] + y, nd->mValue[2] + z)] = 0.0f;
        }
    }
    float r2 = radius * radius;
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), trans, w)) continue;
        float wt = weights ? weights[i] : 1.0f;
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)floorf(w[a] - radius);
            hi[a] = (int64_t)ceilf(w[a] + radius);
            if (lo[a] < 0) lo[a] = 0;
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    float d2 = dist2f((float)x - w[0], (float)y - w[1], (float)z - w[2]);
                    if (!(d2 < r2)) continue;
                    int64_t leaf = voxel_leaf(v, x, y, z);
                    if (leaf < 0) continue;
                    const or_node* nd = &v->nodes[0][leaf];
                    float t = 1.0f - d2 / r2;
                    float k = (t * t) * t;
                    int64_t ai = atlas_index(v, nd->mValue[0] + (x - nd->mPos[0]), nd->mValue[1] + (y - nd->mPos[1]),
                                             nd->mValue[2] + (z - nd->mPos[2]));
                    A[ai] = A[ai] + wt * k;
                }
    }
    return OR_OK;
}

int or_splat_cubic_f64(or_volume* v, const float* pos, int64_t n, int64_t stride, const float* weights,
                       float radius, const float trans[3], double* out) {
    if (!v || (!pos && n > 0) || !out) return OR_ERR_ARG;
    if (stride == 0) stride = 12;
    int r0 = v->res[0];
    int64_t nb = v->ncnt[0];
    memset(out, 0, (size_t)(nb * r0 * r0 * r0) * sizeof(double));
    double R = radius, R2 = R * R;
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), trans, w)) continue;
        double wt = weights ? weights[i] : 1.0;
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)floor((double)w[a] - R);
            hi[a] = (int64_t)ceil((double)w[a] + R);
            if (lo[a] < 0) lo[a] = 0;
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    double dx = x - (double)w[0], dy = y - (double)w[1], dz = z - (double)w[2];
                    double d2 = dx * dx + dy * dy + dz * dz;
                    if (!(d2 < R2)) continue;
                    int64_t leaf = voxel_leaf(v, x, y, z);
                    if (leaf < 0) continue;
                    const or_node* nd = &v->nodes[0][leaf];
                    double t = 1.0 - d2 / R2;
                    int64_t lx = x - nd->mPos[0], ly = y - nd->mPos[1], lz = z - nd->mPos[2];
                    out[leaf * r0 * r0 * r0 + (lz * r0 + ly) * r0 + lx] += wt * t * t * t;
                }
    }
    return OR_OK;
}

int or_fill_channel(or_volume* v, int chan, float val) {
    if (!v || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan]) return OR_ERR_ARG;
    if (v->chan_type[chan] == OR_T_UCHAR4) {
        /* FillChannel VG.cpp:2456-2470: all-zero clears (ClearChannel), else
         * gvdbOpFillC4 OP.cuh:368-373 stores CLR2INT(val, val, val, 1) */
        uint32_t w = 0;
        if (val != 0.0f) {
            uint32_t c = sat_u8(val * 255.0f);
            w = c | c << 8 | c << 16 | 255u << 24;
        }
        uint32_t* A = (uint32_t*)v->atlas[chan];
        for (int64_t i = 0; i < v->atlas_vox; i++) A[i] = w;
        return OR_OK;
    }
    for (int64_t i = 0; i < v->atlas_vox; i++) v->atlas[chan][i] = val;
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* colour channels (uchar4, packed x | y<<8 | z<<16 | w<<24)           */
/* ------------------------------------------------------------------ */

/* Colour mean of GatherDensity / GatherLevelSet (PT.cuh:718-731,757-771)
 * under D8: every interior voxel of a subcell gets
 *   f_a  = init_a + (float)S_a / 255     S_a = sum of component a over the
 *                                        subcell's list (exact integer)
 *   out_a = sat_u8((f_a / (float)(cnt + accum)) * 255)
 * with init = (0,0,0,1), or the stored colour / 255 when accumulating. */
int or_gather_color(or_volume* v, const uint8_t* clr, int64_t clr_stride, int chan, int accum) {
    if (!v || !clr || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan] || v->chan_type[chan] != OR_T_UCHAR4)
        return OR_ERR_ARG;
    if (!v->sc_cnt || !v->sc_nid || !v->sc_pidx) return OR_ERR_STATE;
    if (clr_stride == 0) clr_stride = 4;
    uint32_t* A = (uint32_t*)v->atlas[chan];
    int sc = v->sc_size;
    for (int64_t s = 0; s < v->num_sc; s++) {
        int32_t nid = v->sc_nid[s];
        if (nid < 0) continue;
        const or_node* nd = &v->nodes[0][nid];
        uint64_t S[4] = {0, 0, 0, 0};
        for (int64_t j = v->sc_off[s]; j < v->sc_off[s] + v->sc_cnt[s]; j++) {
            const uint8_t* c = clr + v->sc_pidx[j] * clr_stride;
            for (int a = 0; a < 4; a++) S[a] += c[a];
        }
        const float n = (float)(v->sc_cnt[s] + (accum ? 1 : 0));
        for (int tz = 0; tz < sc; tz++)
            for (int ty = 0; ty < sc; ty++)
                for (int tx = 0; tx < sc; tx++) {
                    int64_t vx = nd->mValue[0] + (v->sc_pos[s * 3] + tx - nd->mPos[0]);
                    int64_t vy = nd->mValue[1] + (v->sc_pos[s * 3 + 1] + ty - nd->mPos[1]);
                    int64_t vz = nd->mValue[2] + (v->sc_pos[s * 3 + 2] + tz - nd->mPos[2]);
                    int64_t ai = atlas_index(v, vx, vy, vz);
                    uint32_t out = 0;
                    for (int a = 0; a < 4; a++) {
                        float init = accum ? (float)((A[ai] >> (8 * a)) & 255u) / 255.0f : (a == 3 ? 1.0f : 0.0f);
                        float f = init + (float)S[a] / 255.0f;
                        out |= sat_u8((f / n) * 255.0f) << (8 * a);
                    }
                    A[ai] = out;
                }
    }
    return OR_OK;
}

/* Colour part of gvdbScatterPointDensity (PT.cuh:554-572) and
 * gvdbScatterPointAvgCol (PT.cuh:610-642): each point colours its home voxel
 * (the voxel the density pass adds to, same validity tests).  avg: per voxel
 * the uint32 sums of x, y, z and the count, then (Sx/cnt, Sy/cnt, Sz/cnt,
 * 255) in integer division where cnt > 0; otherwise the colour of the
 * highest-numbered point (D9). */
int or_scatter_color(or_volume* v, const float* pos, int64_t n, int64_t stride, const uint8_t* clr,
                     int64_t clr_stride, const float trans[3], int avg, int chan) {
    if (!v || (!pos && n > 0) || !clr || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan] ||
        v->chan_type[chan] != OR_T_UCHAR4)
        return OR_ERR_ARG;
    if (stride == 0) stride = 12;
    if (clr_stride == 0) clr_stride = 4;
    uint32_t* A = (uint32_t*)v->atlas[chan];
    int r0 = v->res[0];
    int64_t nvox = v->ncnt[0] * (int64_t)r0 * r0 * r0;
    uint32_t* acc = avg ? (uint32_t*)calloc((size_t)nvox * 4, sizeof(uint32_t)) : NULL;
    for (int64_t i = 0; i < n; i++) {
        float w[3];
        if (!point_valid(pt(pos, i, stride), trans, w)) continue;
        int64_t leaf = point_leaf(v, w);
        if (leaf < 0) continue;
        const or_node* nd = &v->nodes[0][leaf];
        float p[3] = {w[0] - (float)nd->mPos[0], w[1] - (float)nd->mPos[1], w[2] - (float)nd->mPos[2]};
        float pi[3] = {(float)(int)p[0], (float)(int)p[1], (float)(int)p[2]};
        if (pi[0] < 0 || pi[1] < 0 || pi[2] < 0 || pi[0] >= r0 || pi[1] >= r0 || pi[2] >= r0) continue;
        const uint8_t* c = clr + i * clr_stride;
        if (avg) {
            int64_t vid = leaf * r0 * r0 * r0 + ((int64_t)pi[2] * r0 + (int64_t)pi[1]) * r0 + (int64_t)pi[0];
            acc[vid * 4 + 0] += 1u;
            acc[vid * 4 + 1] += c[0];
            acc[vid * 4 + 2] += c[1];
            acc[vid * 4 + 3] += c[2];
        } else {
            int64_t ai = atlas_index(v, (int64_t)pi[0] + nd->mValue[0], (int64_t)pi[1] + nd->mValue[1],
                                     (int64_t)pi[2] + nd->mValue[2]);
            A[ai] = (uint32_t)c[0] | (uint32_t)c[1] << 8 | (uint32_t)c[2] << 16 | (uint32_t)c[3] << 24;
        }
    }
    if (avg) {
        for (int64_t vid = 0; vid < nvox; vid++) {
            uint32_t cnt = acc[vid * 4];
            if (cnt == 0) continue;
            int64_t leaf = vid / ((int64_t)r0 * r0 * r0), l = vid % ((int64_t)r0 * r0 * r0);
            const or_node* nd = &v->nodes[0][leaf];
            int64_t ai = atlas_index(v, nd->mValue[0] + l % r0, nd->mValue[1] + (l / r0) % r0,
                                     nd->mValue[2] + l / ((int64_t)r0 * r0));
            A[ai] = (acc[vid * 4 + 1] / cnt) | (acc[vid * 4 + 2] / cnt) << 8 | (acc[vid * 4 + 3] / cnt) << 16 |
                    255u << 24;
        }
        free(acc);
    }
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* readback                                                            */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* VBX file (SaveVBX VG.cpp:1626-1767, GVDB_FILESPEC.txt)              */
/* ------------------------------------------------------------------ */

static int wr(FILE* fp, const void* p, size_t n) { return fwrite(p, 1, n, fp) == n ? 0 : 1; }

/* SaveVBX, version 1.11: header with the grid transform (pretrans, angles,
 * scale, translation: the defaults (0,0,0), (0,0,0), (1,1,1), (0,0,0)), one
 * grid of float scalars in atlas layout with GVDB topology: per level
 * (log2 dim, res, range xyz, pool 0 count and width, pool 1 count and
 * width), then the node pools and the child-list pools level by level, then
 * per channel its type and stride and the 3-D atlas slice by slice.  Atlas
 * bricks no active leaf owns are written as zeros; the grid name is zeros. */
int or_save_vbx(or_volume* v, const char* path) {
    if (!v || !v->configured || !path) return OR_ERR_ARG;
    if (v->ncnt[0] == 0 || !v->amap) return OR_ERR_STATE;
    int nch = 0;
    while (nch < OR_MAXCHAN && v->has_chan[nch]) nch++;
    for (int c = nch; c < OR_MAXCHAN; c++)
        if (v->has_chan[c]) return OR_ERR_ARG;   /* channels must be 0..n-1 */
    FILE* fp = fopen(path, "wb");
    if (!fp) return OR_ERR_ARG;
    int err = 0;
    const unsigned char major = 1, minor = 11;
    const float xf[12] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0};
    const int num_grids = 1;
    err |= wr(fp, &major, 1) | wr(fp, &minor, 1) | wr(fp, xf, sizeof(xf)) | wr(fp, &num_grids, 4);
    uint64_t grid_off = 2 + sizeof(xf) + 4 + 8;
    err |= wr(fp, &grid_off, 8);
    char name[256];
    memset(name, 0, sizeof(name));
    const char dtype = 'f', comps = 1, compress = 0, topotype = 2, layout = 0;
    const float vsize[3] = {1, 1, 1};
    const int leafcnt = (int)v->ncnt[0];
    const int r0 = v->res[0];
    const int leafdim[3] = {r0, r0, r0};
    const int apron = v->apron;
    const int br = brick_res(v);
    const int axisres[3] = {v->subdim[0] * br, v->subdim[1] * br, v->subdim[2] * br};
    const uint64_t atlas_sz = (uint64_t)4 * axisres[0] * (uint64_t)axisres[1] * axisres[2];
    const int reuse = 0;
    err |= wr(fp, name, 256) | wr(fp, &dtype, 1) | wr(fp, &comps, 1) | wr(fp, &compress, 1) | wr(fp, vsize, 12);
    err |= wr(fp, &leafcnt, 4) | wr(fp, leafdim, 12) | wr(fp, &apron, 4) | wr(fp, &nch, 4) | wr(fp, &atlas_sz, 8);
    err |= wr(fp, &topotype, 1) | wr(fp, &reuse, 4) | wr(fp, &layout, 1) | wr(fp, v->subdim, 12) | wr(fp, axisres, 12);
    const int levels = v->levs;
    const uint64_t root = (uint64_t)v->depth << 8;   /* Elem(0, depth, 0) */
    err |= wr(fp, &levels, 4) | wr(fp, &root, 8);
    for (int l = 0; l < levels; l++) {
        const int ld = v->logdim[l], res = v->res[l];
        const int rg = (int)v->range[l];
        const int range[3] = {rg, rg, rg};
        const int cnt0 = (int)v->ncnt[l], w0 = (int)sizeof(or_node);
        const int cnt1 = l == 0 ? 0 : (int)v->ccnt[l];
        const int w1 = l == 0 ? 0 : (int)(8 * ipow3(res));   /* pool 1 of level 0 is a placeholder (VG.cpp:2427) */
        err |= wr(fp, &ld, 4) | wr(fp, &res, 4) | wr(fp, range, 12) | wr(fp, &cnt0, 4) | wr(fp, &w0, 4);
        err |= wr(fp, &cnt1, 4) | wr(fp, &w1, 4);
    }
    for (int l = 0; l < levels; l++)
        if (v->ncnt[l]) err |= wr(fp, v->nodes[l], (size_t)v->ncnt[l] * sizeof(or_node));
    for (int l = 1; l < levels; l++)
        if (v->ccnt[l]) err |= wr(fp, v->clist[l], (size_t)(v->ccnt[l] * ipow3(v->res[l])) * 8);
    /* atlas: the voxels of every mapped brick, zeros elsewhere */
    const int64_t slice = (int64_t)axisres[0] * axisres[1];
    float* buf = (float*)malloc((size_t)slice * 4);
    for (int c = 0; c < nch; c++) {
        const int type = v->chan_type[c], stride = 4;
        err |= wr(fp, &type, 4) | wr(fp, &stride, 4);
        for (int z = 0; z < axisres[2]; z++) {
            const int bz = z / br;
            memset(buf, 0, (size_t)slice * 4);
            for (int by = 0; by < v->subdim[1]; by++)
                for (int bx = 0; bx < v->subdim[0]; bx++) {
                    const int64_t slot = ((int64_t)bz * v->subdim[1] + by) * v->subdim[0] + bx;
                    if (slot >= v->amap_cnt || v->amap[slot].mLeafID == (int32_t)OR_ID_UNDEFL) continue;
                    for (int y = by * br; y < (by + 1) * br; y++)
                        for (int x = bx * br; x < (bx + 1) * br; x++)
                            buf[(int64_t)y * axisres[0] + x] = v->atlas[c][atlas_index(v, x, y, z)];
                }
            err |= wr(fp, buf, (size_t)slice * 4);
        }
    }
    free(buf);
    if (fclose(fp) != 0) err = 1;
    return err ? OR_ERR_STATE : OR_OK;
}

int or_get_info(or_volume* v, or_info* out) {
    if (!v || !out) return OR_ERR_ARG;
    memset(out, 0, sizeof(*out));
    out->levs = v->levs;
    out->depth = v->depth;
    out->top_lev = v->top_lev;
    memcpy(out->root_pos, v->root_pos, sizeof(v->root_pos));
    for (int l = 0; l < OR_MAXLEV; l++) out->node_cnt[l] = v->ncnt[l];
    memcpy(out->subdim, v->subdim, sizeof(v->subdim));
    out->brick_res = brick_res(v);
    out->apron = v->apron;
    memcpy(out->pos_min, v->pos_min, sizeof(v->pos_min));
    memcpy(out->pos_max, v->pos_max, sizeof(v->pos_max));
    memcpy(out->obj_min, v->obj_min, sizeof(v->obj_min));
    memcpy(out->obj_max, v->obj_max, sizeof(v->obj_max));
    out->sc_len = v->sc_len;
    return OR_OK;
}

const or_node* or_nodes(or_volume* v, int lev) { return (v && lev >= 0 && lev < OR_MAXLEV) ? v->nodes[lev] : NULL; }
const uint64_t* or_child_lists(or_volume* v, int lev) { return (v && lev >= 0 && lev < OR_MAXLEV) ? v->clist[lev] : NULL; }
const or_atlas_node* or_atlas_map(or_volume* v, int64_t* count) {
    if (count) *count = v ? v->amap_cnt : 0;
    return v ? v->amap : NULL;
}
const float* or_atlas(or_volume* v, int chan) { return (v && chan >= 0 && chan < OR_MAXCHAN) ? v->atlas[chan] : NULL; }

/* test input: overwrite one leaf's padded brick (the inverse of or_read_brick) */
int or_write_brick(or_volume* v, int chan, int64_t leaf, const float* in) {
    if (!v || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan] || leaf < 0 || leaf >= v->ncnt[0]) return OR_ERR_ARG;
    int br = brick_res(v);
    const or_node* nd = &v->nodes[0][leaf];
    int64_t c0 = nd->mValue[0] - v->apron, c1 = nd->mValue[1] - v->apron, c2 = nd->mValue[2] - v->apron;
    for (int z = 0; z < br; z++)
        for (int y = 0; y < br; y++)
            for (int x = 0; x < br; x++)
                v->atlas[chan][atlas_index(v, c0 + x, c1 + y, c2 + z)] = in[((int64_t)z * br + y) * br + x];
    return OR_OK;
}

int or_read_brick(or_volume* v, int chan, int64_t leaf, float* out) {
    if (!v || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan] || leaf < 0 || leaf >= v->ncnt[0]) return OR_ERR_ARG;
    int br = brick_res(v);
    const or_node* nd = &v->nodes[0][leaf];
    int64_t c0 = nd->mValue[0] - v->apron, c1 = nd->mValue[1] - v->apron, c2 = nd->mValue[2] - v->apron;
    for (int z = 0; z < br; z++)
        for (int y = 0; y < br; y++)
            for (int x = 0; x < br; x++)
                out[((int64_t)z * br + y) * br + x] = v->atlas[chan][atlas_index(v, c0 + x, c1 + y, c2 + z)];
    return OR_OK;
}


/* ------------------------------------------------------------------ */
/* raycast: HDDA traversal + brick functions                           */
/* ------------------------------------------------------------------ */
/*
 * Every float expression below is written in the reference's operand order
 * (cuda_math.cuh operators, evaluated left to right, no contraction):
 *   dot(a,b) = a.x*b.x + a.y*b.y + a.z*b.z;  normalize(v) = v * (1/sqrtf(dot(v,v)))
 *   (reference: rsqrtf, an approximation; here the correctly rounded form).
 * Trilinear sampling (reference: hardware tex3D with 8-bit weights) is the
 * exact FP32 interpolation of the brick's padded voxels, voxel i centred at
 * i + 0.5 in brick-local coordinates (RC.cuh:104-108 texel convention):
 *   c = p + 0.5, i = floor(c), a = c - i, x then y then z,
 *   lerp(u, w, a) = u*(1-a) + w*a; indices clamped to the padded brick.
 * D7: an HDDA axis whose direction component is zero never steps (tSide =
 *     +inf).  The reference computes mask*tDel = 0*inf = NaN there
 *     (DDA.cuh:60-89), which stalls every ray with a zero component; for all
 *     other rays the two are identical.
 */

typedef struct { float x, y, z; } f3;
static f3 f3m(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 f3add(f3 a, f3 b) { return f3m(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 f3sub(f3 a, f3 b) { return f3m(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 f3scale(f3 a, float s) { return f3m(a.x * s, a.y * s, a.z * s); }
static float f3dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 f3norm(f3 a) { return f3scale(a, 1.0f / sqrtf(f3dot(a, a))); }

enum { M_DEEP_LUT = 0, M_DEEP_LINEAR = 1, M_SURF_TRI = 2, M_SURF_LS = 3, M_SURF_VOXEL = 4, M_SURF_CUBIC = 5,
       M_EMPTY = 6 };

typedef struct {
    int64_t leaf;
    f3 p;
    int lut;
} or_sample;

typedef struct {
    const or_volume* v;
    const float* A;
    const double* B64;
    const or_render_params* rp;
    int mode;
    /* per ray */
    f3 hit, norm;
    float clr[4];
    double clr64[4];
    int64_t hit_leaf;
    f3 hit_p;
    /* colour channel (SetColorChannel): its atlas words, or NULL; the colour
     * the surface brick functions return with a hit (getColorF, 0..255) */
    const uint32_t* C;
    float hclr[4];
    /* optional sample recording (EA adjoint) */
    or_sample* rec;
    int64_t nrec, caprec;
} ray_ctx;

/* getColorF RC.cuh:206-209 with make_float4(uchar4) cuda_math.cuh:221-224:
 * the uchar4 of the colour channel at voxel floor(p) (brick-local p, the
 * interior starting at 0), each byte / 255 in FP32, x from the lowest byte.  The reference truncates p + the brick's atlas
 * offset in FP32; floor(p) is the same voxel except where that sum rounds
 * up to the next integer. */
static void clr_at(const ray_ctx* c, int64_t leaf, f3 p, float out[4]) {
    const or_node* nd = &c->v->nodes[0][leaf];
    int br = brick_res(c->v), a = c->v->apron;
    int q[3] = {(int)floorf(p.x) + a, (int)floorf(p.y) + a, (int)floorf(p.z) + a};
    for (int k = 0; k < 3; k++) q[k] = q[k] < 0 ? 0 : (q[k] >= br ? br - 1 : q[k]);
    uint32_t w = c->C[atlas_index(c->v, (int64_t)nd->mValue[0] - a + q[0], (int64_t)nd->mValue[1] - a + q[1],
                                  (int64_t)nd->mValue[2] - a + q[2])];
    for (int k = 0; k < 4; k++) out[k] = (float)((w >> (8 * k)) & 255u) / 255.0f;
}

static float tri_at32(const ray_ctx* c, const or_node* nd, int i, int j, int k) {
    int br = brick_res(c->v);
    i = i < 0 ? 0 : (i >= br ? br - 1 : i);
    j = j < 0 ? 0 : (j >= br ? br - 1 : j);
    k = k < 0 ? 0 : (k >= br ? br - 1 : k);
    int a = c->v->apron;
    return c->A[atlas_index(c->v, (int64_t)nd->mValue[0] - a + i, (int64_t)nd->mValue[1] - a + j,
                            (int64_t)nd->mValue[2] - a + k)];
}

static double tri_at64(const ray_ctx* c, int64_t leaf, int i, int j, int k) {
    int br = brick_res(c->v);
    i = i < 0 ? 0 : (i >= br ? br - 1 : i);
    j = j < 0 ? 0 : (j >= br ? br - 1 : j);
    k = k < 0 ? 0 : (k >= br ? br - 1 : k);
    if (c->B64) return c->B64[leaf * ipow3(br) + ((int64_t)k * br + j) * br + i];
    return (double)tri_at32(c, &c->v->nodes[0][leaf], i, j, k);
}

typedef struct {
    int i[3];
    float a[3];
} tri_cell;

static tri_cell tri_locate(f3 p) {
    tri_cell t;
    float q[3] = {p.x + 0.5f, p.y + 0.5f, p.z + 0.5f};
    for (int d = 0; d < 3; d++) {
        float f = floorf(q[d]);
        t.i[d] = (int)f;
        t.a[d] = q[d] - f;
    }
    return t;
}

static float tri32(const ray_ctx* c, int64_t leaf, f3 p) {
    const or_node* nd = &c->v->nodes[0][leaf];
    tri_cell t = tri_locate(p);
    int ix = t.i[0], iy = t.i[1], iz = t.i[2];
    float ax = t.a[0], ay = t.a[1], az = t.a[2];
    float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    float c00 = tri_at32(c, nd, ix, iy, iz) * bx + tri_at32(c, nd, ix + 1, iy, iz) * ax;
    float c10 = tri_at32(c, nd, ix, iy + 1, iz) * bx + tri_at32(c, nd, ix + 1, iy + 1, iz) * ax;
    float c01 = tri_at32(c, nd, ix, iy, iz + 1) * bx + tri_at32(c, nd, ix + 1, iy, iz + 1) * ax;
    float c11 = tri_at32(c, nd, ix, iy + 1, iz + 1) * bx + tri_at32(c, nd, ix + 1, iy + 1, iz + 1) * ax;
    float c0 = c00 * by + c10 * ay;
    float c1 = c01 * by + c11 * ay;
    return c0 * bz + c1 * az;
}

/* weights of the 8 corners, corner index (dz*2+dy)*2+dx */
static void tri_weights64(const tri_cell* t, double w[8]) {
    double a[3] = {t->a[0], t->a[1], t->a[2]};
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++)
            for (int dx = 0; dx < 2; dx++)
                w[(dz * 2 + dy) * 2 + dx] = (dx ? a[0] : 1.0 - a[0]) * (dy ? a[1] : 1.0 - a[1]) * (dz ? a[2] : 1.0 - a[2]);
}

static double tri64(const ray_ctx* c, int64_t leaf, f3 p) {
    tri_cell t = tri_locate(p);
    double w[8], s = 0.0;
    tri_weights64(&t, w);
    for (int k = 0; k < 8; k++)
        s += w[k] * tri_at64(c, leaf, t.i[0] + (k & 1), t.i[1] + ((k >> 1) & 1), t.i[2] + (k >> 2));
    return s;
}

/* rayBoxIntersect GEO.cuh:85-98 */
static f3 ray_box(f3 pos, f3 dir, f3 vmin, f3 vmax) {
    float ht[8];
    ht[0] = (vmin.x - pos.x) / dir.x;
    ht[1] = (vmax.x - pos.x) / dir.x;
    ht[2] = (vmin.y - pos.y) / dir.y;
    ht[3] = (vmax.y - pos.y) / dir.y;
    ht[4] = (vmin.z - pos.z) / dir.z;
    ht[5] = (vmax.z - pos.z) / dir.z;
    ht[6] = fmaxf(fmaxf(fminf(ht[0], ht[1]), fminf(ht[2], ht[3])), fminf(ht[4], ht[5]));
    ht[7] = fminf(fminf(fmaxf(ht[0], ht[1]), fmaxf(ht[2], ht[3])), fmaxf(ht[4], ht[5]));
    ht[6] = (ht[6] < 0) ? 0.0f : ht[6];
    return f3m(ht[6], ht[7], (ht[7] < ht[6] || ht[7] < 0) ? OR_NOHIT : 0.0f);
}

/* HDDAState DDA.cuh:38-91 (+ D7) */
typedef struct {
    f3 pos, dir;
    int ps[3];
    float tdel[3];
    f3 t;
    int p[3];
    float ts[3];
    int mask[3];
} hdda;

static void hdda_set(hdda* d, f3 pos, f3 dir, f3 t) {
    d->pos = pos;
    d->dir = dir;
    d->ps[0] = dir.x > 0 ? 1 : -1;   /* isign3 cuda_math.cuh:1541-1545 */
    d->ps[1] = dir.y > 0 ? 1 : -1;
    d->ps[2] = dir.z > 0 ? 1 : -1;
    d->t = t;
}

static void hdda_prepare(hdda* d, f3 vmin, float vdel) {
    float dir[3] = {d->dir.x, d->dir.y, d->dir.z};
    f3 q = f3sub(f3add(d->pos, f3scale(d->dir, d->t.x)), vmin);
    float pf[3] = {q.x / vdel, q.y / vdel, q.z / vdel};
    for (int a = 0; a < 3; a++) {
        d->tdel[a] = fabsf(vdel / dir[a]);
        float fl = floorf(pf[a]);
        float br = ((fl - pf[a]) + 0.5f) * (float)d->ps[a] + 0.5f;
        d->ts[a] = isinf(d->tdel[a]) ? INFINITY : br * d->tdel[a] + d->t.x;
        d->p[a] = (int)fl;
    }
}

/* PrepareLeaf DDA.cuh:67-73: vdel 1, and tSide without the "+ t.x" of
 * Prepare (only the order of the steps matters inside a brick) */
static void hdda_prepare_leaf(hdda* d, f3 vmin) {
    float dir[3] = {d->dir.x, d->dir.y, d->dir.z};
    f3 q = f3sub(f3add(d->pos, f3scale(d->dir, d->t.x)), vmin);
    float pf[3] = {q.x, q.y, q.z};
    for (int a = 0; a < 3; a++) {
        d->tdel[a] = fabsf(1.0f / dir[a]);
        float fl = floorf(pf[a]);
        float br = ((fl - pf[a]) + 0.5f) * (float)d->ps[a] + 0.5f;
        d->ts[a] = isinf(d->tdel[a]) ? INFINITY : br * d->tdel[a];
        d->p[a] = (int)fl;
    }
}

static void hdda_next(hdda* d) {
    float* s = d->ts;
    d->mask[0] = (s[0] < s[1]) & (s[0] <= s[2]);
    d->mask[1] = (s[1] < s[2]) & (s[1] <= s[0]);
    d->mask[2] = (s[2] < s[0]) & (s[2] <= s[1]);
    d->t.y = d->mask[0] ? s[0] : (d->mask[1] ? s[1] : s[2]);
}

static void hdda_step(hdda* d) {
    d->t.x = d->t.y;
    for (int a = 0; a < 3; a++) {
        if (d->mask[a]) d->ts[a] = d->ts[a] + d->tdel[a];
        d->p[a] += d->mask[a] * d->ps[a];
    }
}

static void record_sample(ray_ctx* c, int64_t leaf, f3 p, int lut) {
    if (!c->rec && c->caprec == 0) return;
    if (c->nrec >= c->caprec) {
        int64_t nc = c->caprec ? c->caprec * 2 : 256;
        c->rec = (or_sample*)realloc(c->rec, (size_t)nc * sizeof(or_sample));
        c->caprec = nc;
    }
    or_sample s = {leaf, p, lut};
    c->rec[c->nrec++] = s;
}

/* transfer DDA.cuh:20-23: float division, double clamp and scale, truncation */
static int transfer_index(const or_render_params* rp, float s) {
    float q = (s - rp->thresh) / (rp->vmax - rp->vmin);
    double qc = fmin(1.0, fmax(0.0, (double)q));
    return (int)(qc * 16300.0f);
}

/* rayDeepBrick RC.cuh:485-533 (no depth buffer).
 * M_DEEP_LUT uses the transfer function; M_DEEP_LINEAR is the transfer
 * s -> (s,s,s,s) with no final clamp (the differentiable form). */
static void brick_deep(ray_ctx* c, int64_t leaf, f3 t, f3 pos, f3 dir) {
    const or_volume* v = c->v;
    const or_render_params* rp = c->rp;
    const or_node* nd = &v->nodes[0][leaf];
    f3 vmin = f3m((float)nd->mPos[0], (float)nd->mPos[1], (float)nd->mPos[2]);
    float step = rp->step;
    t.x = step * ceilf(t.x / step);
    f3 wp = f3add(pos, f3scale(dir, t.x));
    f3 p = f3sub(wp, vmin);
    f3 wpt = f3scale(dir, step);
    float dt = sqrtf(f3dot(wpt, wpt));
    float r0 = (float)v->res[0];
    if (c->hit.x == 0) c->hit.x = t.x;
    for (int iter = 0; c->clr[3] > rp->alphacut && iter < rp->max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 &&
                       p.x < r0 && p.y < r0 && p.z < r0;
         iter++) {
        float s = tri32(c, leaf, p);
        if (s >= rp->minval) {
            float val[4];
            double v64[4];
            int li = -1;
            if (c->mode == M_DEEP_LUT) {
                li = transfer_index(rp, s);
                for (int k = 0; k < 4; k++) val[k] = rp->lut[4 * li + k];
                for (int k = 0; k < 4; k++) v64[k] = val[k];
            } else {
                double s64 = tri64(c, leaf, p);
                for (int k = 0; k < 4; k++) { val[k] = s; v64[k] = s64; }
            }
            record_sample(c, leaf, p, li);
            /* RGB from the colour channel (RC.cuh:519), else 1 */
            float h[4] = {1, 1, 1, 1};
            if (c->C && c->mode == M_DEEP_LUT) clr_at(c, leaf, p, h);
            val[3] = expf(rp->extinct * val[3] * step);
            for (int k = 0; k < 3; k++) c->clr[k] += val[k] * c->clr[3] * (1 - val[3]) * rp->albedo * h[k];
            c->clr[3] *= val[3];
            double a64 = exp((double)rp->extinct * v64[3] * (double)step);
            for (int k = 0; k < 3; k++)
                c->clr64[k] += v64[k] * c->clr64[3] * (1.0 - a64) * (double)rp->albedo * (double)h[k];
            c->clr64[3] *= a64;
        }
        p = f3add(p, wpt);
        wp = f3add(wp, wpt);
        t.x += dt;
    }
    c->hit.y = t.x;
    if (c->mode == M_DEEP_LUT) {
        for (int k = 0; k < 3; k++) {
            c->clr[k] = fminf(c->clr[k], 1.f);
            c->clr64[k] = fmin(c->clr64[k], 1.0);
        }
        c->clr[3] = fmaxf(c->clr[3], 0.f);
        c->clr64[3] = fmax(c->clr64[3], 0.0);
    }
}

/* getGradient RC.cuh:132-142 (sign -1) and getGradientLevelSet RC.cuh:146-157
 * (sign +1), unnormalised, FP32 */
static f3 gradient32(const ray_ctx* c, int64_t leaf, f3 p, int sign) {
    f3 g;
    float lo, hi;
    lo = tri32(c, leaf, f3m(p.x - 0.5f, p.y, p.z)); hi = tri32(c, leaf, f3m(p.x + 0.5f, p.y, p.z));
    g.x = sign > 0 ? hi - lo : lo - hi;
    lo = tri32(c, leaf, f3m(p.x, p.y - 0.5f, p.z)); hi = tri32(c, leaf, f3m(p.x, p.y + 0.5f, p.z));
    g.y = sign > 0 ? hi - lo : lo - hi;
    lo = tri32(c, leaf, f3m(p.x, p.y, p.z - 0.5f)); hi = tri32(c, leaf, f3m(p.x, p.y, p.z + 0.5f));
    g.z = sign > 0 ? hi - lo : lo - hi;
    return f3norm(g);
}

static void gradient64(const ray_ctx* c, int64_t leaf, f3 p, int sign, double g[3]) {
    double lo, hi;
    lo = tri64(c, leaf, f3m(p.x - 0.5f, p.y, p.z)); hi = tri64(c, leaf, f3m(p.x + 0.5f, p.y, p.z));
    g[0] = sign > 0 ? hi - lo : lo - hi;
    lo = tri64(c, leaf, f3m(p.x, p.y - 0.5f, p.z)); hi = tri64(c, leaf, f3m(p.x, p.y + 0.5f, p.z));
    g[1] = sign > 0 ? hi - lo : lo - hi;
    lo = tri64(c, leaf, f3m(p.x, p.y, p.z - 0.5f)); hi = tri64(c, leaf, f3m(p.x, p.y, p.z + 0.5f));
    g[2] = sign > 0 ? hi - lo : lo - hi;
    double n = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    for (int k = 0; k < 3; k++) g[k] /= n;
}

/* raySurfaceTrilinearBrick RC.cuh:281-300 */
static void brick_surf_tri(ray_ctx* c, int64_t leaf, f3 t, f3 pos, f3 dir) {
    const or_volume* v = c->v;
    const or_render_params* rp = c->rp;
    const or_node* nd = &v->nodes[0][leaf];
    f3 vmin = f3m((float)nd->mPos[0], (float)nd->mPos[1], (float)nd->mPos[2]);
    t.x = rp->step * ceilf(t.x / rp->step);
    f3 p = f3sub(f3add(pos, f3scale(dir, t.x)), vmin);
    float r0 = (float)v->res[0];
    for (int iter = 0; iter < rp->max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0;
         iter++) {
        if (tri32(c, leaf, p) >= rp->thresh) {
            c->hit = f3add(p, vmin);
            c->norm = gradient32(c, leaf, p, -1);
            c->hit_leaf = leaf;
            c->hit_p = p;
            if (c->C) clr_at(c, leaf, p, c->hclr);
            return;
        }
        p = f3add(p, f3scale(dir, rp->step));
        t.x += rp->step;
    }
}

/* rayLevelSetBrick RC.cuh:389-410; its rayLevelSet (RC.cuh:186-197) repeats
 * the caller's test at the same point and so returns p + vmin at once. */
static void brick_surf_ls(ray_ctx* c, int64_t leaf, f3 t, f3 pos, f3 dir) {
    const or_volume* v = c->v;
    const or_render_params* rp = c->rp;
    const or_node* nd = &v->nodes[0][leaf];
    f3 vmin = f3m((float)nd->mPos[0], (float)nd->mPos[1], (float)nd->mPos[2]);
    f3 p = f3sub(f3add(pos, f3scale(dir, t.x)), vmin);
    float r0 = (float)v->res[0];
    for (int iter = 0; iter < rp->max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x <= r0 && p.y <= r0 && p.z <= r0;
         iter++) {
        if (tri32(c, leaf, p) < rp->thresh) {
            c->hit = f3add(p, vmin);
            c->norm = gradient32(c, leaf, p, +1);
            c->hit_leaf = leaf;
            c->hit_p = p;
            if (c->C) clr_at(c, leaf, p, c->hclr);
            return;
        }
        p = f3add(p, f3scale(dir, rp->step));
    }
}

/* raySurfaceVoxelBrick RC.cuh:227-263: the leaf's voxels in DDA order, the
 * first one above the threshold is hit as a unit cube (its exact value: the
 * reference samples the texel centre, p + o + 0.5); a missed cube repeats
 * the same voxel (no step) until the iteration limit, as the reference's
 * `continue` does. */
static void brick_surf_voxel(ray_ctx* c, int64_t leaf, f3 t, f3 pos, f3 dir) {
    const or_volume* v = c->v;
    const or_render_params* rp = c->rp;
    const or_node* nd = &v->nodes[0][leaf];
    f3 vmin = f3m((float)nd->mPos[0], (float)nd->mPos[1], (float)nd->mPos[2]);
    int r0 = v->res[0], a = v->apron;
    hdda d;
    hdda_set(&d, pos, dir, t);
    hdda_prepare_leaf(&d, vmin);
    for (int iter = 0; iter < rp->max_iter && d.p[0] >= 0 && d.p[1] >= 0 && d.p[2] >= 0 && d.p[0] < r0 &&
                       d.p[1] < r0 && d.p[2] < r0;
         iter++) {
        if (tri_at32(c, nd, d.p[0] + a, d.p[1] + a, d.p[2] + a) > rp->thresh) {
            vmin = f3add(vmin, f3m((float)d.p[0], (float)d.p[1], (float)d.p[2]));
            f3 tb = ray_box(pos, dir, vmin, f3add(vmin, f3m(1.0f, 1.0f, 1.0f)));
            d.t = tb;
            if (tb.z == OR_NOHIT) {
                c->hit.z = OR_NOHIT;
                continue;
            }
            f3 hit = f3add(pos, f3scale(dir, tb.x));
            f3 fc = f3sub(f3sub(hit, vmin), f3m(0.5f, 0.5f, 0.5f));
            fc = f3sub(fc, f3scale(dir, 0.01f));
            float m = fmaxf(fmaxf(fabsf(fc.x), fabsf(fc.y)), fabsf(fc.z));
            c->norm = f3m(fabsf(fc.x) == m ? copysignf(1.0f, fc.x) : 0.0f, fabsf(fc.y) == m ? copysignf(1.0f, fc.y) : 0.0f,
                          fabsf(fc.z) == m ? copysignf(1.0f, fc.z) : 0.0f);
            c->hit = hit;
            c->hit_leaf = leaf;
            c->hit_p = f3sub(hit, vmin);
            if (c->C) clr_at(c, leaf, f3m((float)d.p[0], (float)d.p[1], (float)d.p[2]), c->hclr);
            return;
        }
        hdda_next(&d);
        hdda_step(&d);
    }
}

/* getTricubic RC.cuh:32-98: 27 filtered taps at the integer points
 * floor(p) - 1 + {0, 1, 2} per axis (each the trilinear sample there, i.e.
 * the mean of 8 voxels), combined with the weights ta^2, 2 ta tb, tb^2,
 * tb = frac(p)/2 + 1/4, x then y then z.  Brick-local coordinates: the
 * reference forms floor(p + atlas offset) in FP32 and reads past the padded
 * brick into its atlas neighbour, where the taps here clamp to the brick. */
static float tricubic32(const ray_ctx* c, int64_t leaf, f3 p) {
    float fl[3] = {floorf(p.x), floorf(p.y), floorf(p.z)};
    float pv[3] = {p.x, p.y, p.z};
    float ta2[3], tab[3], tb2[3];
    for (int k = 0; k < 3; k++) {
        float tb = (pv[k] - fl[k]) * 0.5f + 0.25f;
        float ta = 1.0f - tb;
        ta2[k] = ta * ta;
        tb2[k] = tb * tb;
        tab[k] = ta * tb * 2.0f;
    }
    float q[3] = {fl[0] - 1.0f, fl[1] - 1.0f, fl[2] - 1.0f};
    float row[3][3];   /* [z][y]: the x-combined value of each (y, z) row */
    for (int zz = 0; zz < 3; zz++)
        for (int yy = 0; yy < 3; yy++) {
            float t0 = tri32(c, leaf, f3m(q[0], q[1] + (float)yy, q[2] + (float)zz));
            float t1 = tri32(c, leaf, f3m(q[0] + 1.0f, q[1] + (float)yy, q[2] + (float)zz));
            float t2 = tri32(c, leaf, f3m(q[0] + 2.0f, q[1] + (float)yy, q[2] + (float)zz));
            row[zz][yy] = t0 * ta2[0] + t1 * tab[0] + t2 * tb2[0];
        }
    float jkl[3];
    for (int zz = 0; zz < 3; zz++) jkl[zz] = row[zz][0] * ta2[1] + row[zz][1] * tab[1] + row[zz][2] * tb2[1];
    return jkl[0] * ta2[2] + jkl[1] * tab[2] + jkl[2] * tb2[2];
}

/* getGradientTricubic RC.cuh:161-171 (the negated gradient, half-voxel
 * central differences) */
static f3 gradient_cubic32(const ray_ctx* c, int64_t leaf, f3 p) {
    const float vs = 0.5f;
    f3 g;
    g.x = (tricubic32(c, leaf, f3add(p, f3m(-vs, 0, 0))) - tricubic32(c, leaf, f3add(p, f3m(vs, 0, 0)))) / (2 * vs);
    g.y = (tricubic32(c, leaf, f3add(p, f3m(0, -vs, 0))) - tricubic32(c, leaf, f3add(p, f3m(0, vs, 0)))) / (2 * vs);
    g.z = (tricubic32(c, leaf, f3add(p, f3m(0, 0, -vs))) - tricubic32(c, leaf, f3add(p, f3m(0, 0, vs)))) / (2 * vs);
    return f3norm(g);
}

/* raySurfaceTricubicBrick RC.cuh:316-340: march from the entry point (no
 * snap to the step grid), at the first sample >= thresh step back by the
 * linear estimate of the crossing over one fine step */
static void brick_surf_cubic(ray_ctx* c, int64_t leaf, f3 t, f3 pos, f3 dir) {
    const or_volume* v = c->v;
    const or_render_params* rp = c->rp;
    const or_node* nd = &v->nodes[0][leaf];
    f3 vmin = f3m((float)nd->mPos[0], (float)nd->mPos[1], (float)nd->mPos[2]);
    f3 p = f3sub(f3add(pos, f3scale(dir, t.x)), vmin);
    float r0 = (float)v->res[0];
    for (int iter = 0; iter < rp->max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0;
         iter++) {
        float vz = tricubic32(c, leaf, p);
        if (vz >= rp->thresh) {
            float vx = tricubic32(c, leaf, f3sub(p, f3scale(dir, rp->finestep)));
            float vy = (vz - rp->thresh) / (vz - vx);
            p = f3add(p, f3scale(dir, -vy * rp->finestep));
            c->hit = f3add(p, vmin);
            c->norm = gradient_cubic32(c, leaf, p);
            c->hit_leaf = leaf;
            c->hit_p = p;
            if (c->C) clr_at(c, leaf, p, c->hclr);
            return;
        }
        p = f3add(p, f3scale(dir, rp->step));
        t.x += rp->step;
    }
}

static void brick_call(ray_ctx* c, int64_t leaf, f3 t, f3 pos, f3 dir) {
    switch (c->mode) {
        case M_DEEP_LUT:
        case M_DEEP_LINEAR: brick_deep(c, leaf, t, pos, dir); break;
        case M_SURF_TRI: brick_surf_tri(c, leaf, t, pos, dir); break;
        case M_SURF_VOXEL: brick_surf_voxel(c, leaf, t, pos, dir); break;
        case M_SURF_CUBIC: brick_surf_cubic(c, leaf, t, pos, dir); break;
        case M_EMPTY: c->hit = f3add(pos, f3scale(dir, t.x)); break;   /* rayEmptySkipBrick RC.cuh:425-428 */
        default: brick_surf_ls(c, leaf, t, pos, dir); break;
    }
}

/* rayCast RC.cuh:543-611 (no depth buffer) */
static void ray_cast(ray_ctx* c, f3 pos, f3 dir) {
    const or_volume* v = c->v;
    const or_render_params* rp = c->rp;
    int top = v->top_lev;
    if (top < 1 || top >= v->levs || v->ncnt[top] < 1) return;
    int64_t nodeid[OR_MAXLEV];
    float tmax[OR_MAXLEV + 1];
    int lev = top;
    nodeid[lev] = 0;
    f3 bmin = f3m(v->obj_min[0], v->obj_min[1], v->obj_min[2]);
    f3 bmax = f3m(v->obj_max[0], v->obj_max[1], v->obj_max[2]);
    f3 ts = ray_box(pos, dir, bmin, bmax);
    const or_node* node = &v->nodes[lev][0];
    f3 vmin = f3m((float)node->mPos[0], (float)node->mPos[1], (float)node->mPos[2]);
    if (ts.z == OR_NOHIT) return;
    ts.x += rp->epsilon;
    tmax[lev] = ts.y - rp->epsilon;
    hdda d;
    hdda_set(&d, pos, dir, ts);
    hdda_prepare(&d, vmin, (float)or_range(v, lev - 1));
    for (int iter = 0; iter < rp->max_iter && lev > 0 && lev <= top && d.p[0] >= 0 && d.p[1] >= 0 && d.p[2] >= 0 &&
                       d.p[0] <= v->res[lev] && d.p[1] <= v->res[lev] && d.p[2] <= v->res[lev];
         iter++) {
        hdda_next(&d);
        int dim = v->logdim[lev];
        int64_t b = ((((int64_t)d.p[2] << dim) + d.p[1]) << dim) + d.p[0];
        int64_t child = get_child(v, node, b);
        if (child >= 0 && child < v->ncnt[lev - 1]) {
            if (lev == 1) {
                d.t.x += rp->epsilon;
                brick_call(c, child, d.t, pos, dir);
                if (c->clr[3] <= 0) {
                    c->clr[3] = 0;
                    return;
                }
                if (c->hit.z != OR_NOHIT) return;
                hdda_step(&d);
            } else {
                lev--;
                nodeid[lev] = child;
                node = &v->nodes[lev][child];
                vmin = f3m((float)node->mPos[0], (float)node->mPos[1], (float)node->mPos[2]);
                d.t.x += rp->epsilon;
                tmax[lev] = d.t.y - rp->epsilon;
                hdda_prepare(&d, vmin, (float)or_range(v, lev - 1));
            }
        } else {
            hdda_step(&d);
        }
        while (lev <= top && d.t.x > tmax[lev]) {
            lev++;
            if (lev <= top) {
                node = &v->nodes[lev][nodeid[lev]];
                vmin = f3m((float)node->mPos[0], (float)node->mPos[1], (float)node->mPos[2]);
                hdda_prepare(&d, vmin, (float)or_range(v, lev - 1));
            }
        }
    }
}

/* pixel ray, gvdbRayDeep MOD.cu:66-71 with GEO.cuh:48-63 (identity transform) */
static void camera_ray(const or_camera* cam, int x, int y, f3* pos, f3* dir) {
    float u = ((float)x + 0.5f) / (float)cam->width;
    float w = ((float)y + 0.5f) / (float)cam->height;
    *pos = f3m((u * cam->org_u[0] + w * cam->org_v[0]) + cam->org[0], (u * cam->org_u[1] + w * cam->org_v[1]) + cam->org[1],
               (u * cam->org_u[2] + w * cam->org_v[2]) + cam->org[2]);
    f3 d = f3m((u * cam->dir_u[0] + w * cam->dir_v[0]) + cam->dir[0], (u * cam->dir_u[1] + w * cam->dir_v[1]) + cam->dir[1],
               (u * cam->dir_u[2] + w * cam->dir_v[2]) + cam->dir[2]);
    *dir = f3norm(d);
}

static int render_check(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp) {
    if (!v || !cam || !rp || chan < 0 || chan >= OR_MAXCHAN || !v->has_chan[chan]) return OR_ERR_ARG;
    if (rp->use_clr && (rp->clr_chan < 0 || rp->clr_chan >= OR_MAXCHAN || !v->has_chan[rp->clr_chan] ||
                        v->chan_type[rp->clr_chan] != OR_T_UCHAR4))
        return OR_ERR_ARG;
    if (cam->width <= 0 || cam->height <= 0 || !(rp->step > 0)) return OR_ERR_ARG;
    if (!v->amap || v->apron != 1) return OR_ERR_STATE;
    return OR_OK;
}

/* performPhongShading MOD.cu:38-57: the hit's colour (1,1,1 without a
 * colour channel) times diffuse + ambient; FP32 and FP64 */
static void phong(ray_ctx* c, f3 pos_unused, float out[4], double out64[4]) {
    (void)pos_unused;
    const or_render_params* rp = c->rp;
    if (c->hit.z == OR_NOHIT) {
        for (int k = 0; k < 4; k++) { out[k] = rp->backclr[k]; out64[k] = rp->backclr[k]; }
        return;
    }
    f3 light = f3m(rp->light[0], rp->light[1], rp->light[2]);
    f3 shit = c->hit, snorm = c->norm;
    int64_t hl = c->hit_leaf;
    f3 hp = c->hit_p;
    f3 ldir = f3norm(f3sub(light, shit));
    float diff = (float)(0.9 * (double)fmaxf(0.0f, f3dot(snorm, ldir)));
    float amb = 0.1f;
    /* FP64 shading at the same hit point (trilinear / level-set normals; the
     * other surfaces repeat the FP32 value) */
    const int fp64 = c->mode == M_SURF_TRI || c->mode == M_SURF_LS;
    double diff64 = diff;
    if (fp64) {
        double g[3];
        gradient64(c, hl, hp, c->mode == M_SURF_LS ? +1 : -1, g);
        double L[3] = {(double)rp->light[0] - shit.x, (double)rp->light[1] - shit.y, (double)rp->light[2] - shit.z};
        double ln = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
        double dd = (g[0] * L[0] + g[1] * L[1] + g[2] * L[2]) / ln;
        diff64 = 0.9 * (dd > 0 ? dd : 0.0);
    }
    if (rp->shadow_amt > 0) {
        ray_ctx s2 = *c;
        /* gvdbRaySurfaceTricubic shades with the trilinear brick (MOD.cu:137) */
        if (c->mode == M_SURF_CUBIC) s2.mode = M_SURF_TRI;
        s2.hit = f3m(0, 0, OR_NOHIT);
        s2.clr[0] = s2.clr[1] = s2.clr[2] = 0; s2.clr[3] = 1;
        s2.rec = NULL; s2.caprec = 0; s2.nrec = 0;
        f3 o = f3add(shit, f3scale(snorm, rp->shadow_bias));
        ray_cast(&s2, o, ldir);
        if (s2.hit.z != OR_NOHIT) {
            diff = (float)((double)diff * (1.0 - (double)rp->shadow_amt));
            diff64 = diff64 * (1.0 - (double)rp->shadow_amt);
        }
    }
    float sh = diff + amb;
    for (int k = 0; k < 3; k++) {
        out[k] = c->hclr[k] * sh;
        out64[k] = (double)c->hclr[k] * (fp64 ? diff64 + 0.1 : (double)sh);
    }
    out[3] = 1.0f;
    out64[3] = 1.0;
}

/* transfer(gvdb, v) with the LUT of the sections: entry transfer_index(v) */
static void lut_at(const or_render_params* rp, float s, float out[4]) {
    int i = transfer_index(rp, s);
    for (int k = 0; k < 4; k++) out[k] = rp->lut[4 * i + k];
}

/* gvdbSection2D MOD.cu:268-293: pixel (x, y) samples the slice point
 * pnt + ((2x/w - 1, 0, 2y/h - 1) * norm) (componentwise, the offset formed
 * in double), black outside every leaf; colour = transfer colour * alpha */
static void section2d(ray_ctx* c, const or_camera* cam, int x, int y, float o[4]) {
    const or_render_params* rp = c->rp;
    float sx = (float)((double)(float)x * 2.0 / (double)cam->width - 1.0);
    float sz = (float)((double)(float)y * 2.0 / (double)cam->height - 1.0);
    float w[3] = {rp->slice_pnt[0] + sx * rp->slice_norm[0], rp->slice_pnt[1] + 0.0f * rp->slice_norm[1],
                  rp->slice_pnt[2] + sz * rp->slice_norm[2]};
    int64_t leaf = point_leaf(c->v, w);
    o[0] = o[1] = o[2] = 0.0f;
    o[3] = 1.0f;
    if (leaf < 0) return;
    const or_node* nd = &c->v->nodes[0][leaf];
    float s = tri32(c, leaf, f3m(w[0] - (float)nd->mPos[0], w[1] - (float)nd->mPos[1], w[2] - (float)nd->mPos[2]));
    float clr[4];
    lut_at(rp, s, clr);
    for (int k = 0; k < 3; k++) o[k] = 0.0f + clr[3] * (clr[k] - 0.0f);
}

/* gvdbSection3D MOD.cu:222-265: the transfer colour where the ray meets the
 * slice plane, blended over the trilinear surface found beyond it */
static void section3d(ray_ctx* c, f3 pos, f3 dir, float o[4]) {
    const or_render_params* rp = c->rp;
    float clr[4] = {1, 1, 1, 0};
    f3 wpos = pos;
    const float* n = rp->slice_norm;
    const float* q = rp->slice_pnt;
    /* rayPlaneIntersect GEO.cuh:67-71 */
    float t = ((q[0] - wpos.x) * n[0] + (q[1] - wpos.y) * n[1] + (q[2] - wpos.z) * n[2]) /
              (dir.x * n[0] + dir.y * n[1] + dir.z * n[2]);
    t = t > 0 ? t : OR_NOHIT;
    if (t > 0) {
        wpos = f3add(wpos, f3scale(dir, t));
        float w[3] = {wpos.x, wpos.y, wpos.z};
        int64_t leaf = point_leaf(c->v, w);
        if (leaf >= 0) {
            const or_node* nd = &c->v->nodes[0][leaf];
            f3 pl = f3m(wpos.x - (float)nd->mPos[0], wpos.y - (float)nd->mPos[1], wpos.z - (float)nd->mPos[2]);
            t = tri32(c, leaf, pl);
            lut_at(rp, t, clr);
            if (c->C) {   /* MOD.cu:249-252: times the colour there (alpha times 1) */
                float h[4];
                clr_at(c, leaf, pl, h);
                for (int k = 0; k < 3; k++) clr[k] *= h[k];
            }
        } else {
            t = 0;
        }
    }
    ray_ctx s = *c;
    s.mode = M_SURF_TRI;
    s.hit = f3m(OR_NOHIT, OR_NOHIT, OR_NOHIT);
    s.norm = f3m(0, 0, 0);
    s.clr[0] = s.clr[1] = s.clr[2] = s.clr[3] = 1;
    for (int k = 0; k < 4; k++) s.hclr[k] = 1.0f;
    ray_cast(&s, wpos, dir);
    if (s.hit.z != OR_NOHIT) {
        f3 ldir = f3norm(f3sub(f3m(rp->light[0], rp->light[1], rp->light[2]), s.hit));
        float ds = (t > rp->thresh) ? 1.0f : (float)(0.8 * (double)fmaxf(0.0f, f3dot(s.norm, ldir)));
        for (int k = 0; k < 4; k++) {
            float a = s.hclr[k] * ds;   /* the surface colour (1 without a colour channel) */
            o[k] = a + clr[3] * (clr[k] - a);
        }
    } else {
        for (int k = 0; k < 4; k++) o[k] = rp->backclr[k] + clr[3] * (clr[k] - rp->backclr[k]);
    }
}

int or_render(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp, const double* bricks64,
              float* out32, double* out64) {
    int st = render_check(v, chan, cam, rp);
    if (st) return st;
    int mode;
    switch (rp->shade) {
        case OR_SHADE_VOXEL: mode = M_SURF_VOXEL; break;
        case OR_SHADE_TRILINEAR: mode = M_SURF_TRI; break;
        case OR_SHADE_TRICUBIC: mode = M_SURF_CUBIC; if (!(rp->finestep > 0)) return OR_ERR_ARG; break;
        case OR_SHADE_LEVELSET: mode = M_SURF_LS; break;
        case OR_SHADE_EMPTYSKIP: mode = M_EMPTY; break;
        case OR_SHADE_SECTION2D:
        case OR_SHADE_SECTION3D: mode = M_SURF_TRI; if (!rp->lut) return OR_ERR_ARG; break;
        case OR_SHADE_VOLUME: mode = M_DEEP_LUT; if (!rp->lut) return OR_ERR_ARG; break;
        case OR_SHADE_OFF:   /* Render VG.cpp:4340-4345: the buffer is cleared */
            for (int64_t i = 0; i < (int64_t)cam->width * cam->height * 4; i++) {
                if (out32) out32[i] = 0.0f;
                if (out64) out64[i] = 0.0;
            }
            return OR_OK;
        default: return OR_ERR_ARG;
    }
    for (int y = 0; y < cam->height; y++)
        for (int x = 0; x < cam->width; x++) {
            ray_ctx c;
            memset(&c, 0, sizeof(c));
            c.v = v; c.A = v->atlas[chan]; c.B64 = bricks64; c.rp = rp; c.mode = mode;
            c.C = rp->use_clr ? (const uint32_t*)v->atlas[rp->clr_chan] : NULL;
            for (int k = 0; k < 4; k++) c.hclr[k] = 1.0f;
            f3 pos, dir;
            camera_ray(cam, x, y, &pos, &dir);
            float o32[4];
            double o64[4];
            if (rp->shade == OR_SHADE_SECTION2D) {
                section2d(&c, cam, x, y, o32);
                for (int k = 0; k < 4; k++) o64[k] = o32[k];
            } else if (rp->shade == OR_SHADE_SECTION3D) {
                section3d(&c, pos, dir, o32);
                for (int k = 0; k < 4; k++) o64[k] = o32[k];
            } else if (mode == M_EMPTY) {
                /* gvdbRayEmptySkip MOD.cu:184-204: the first brick entry point / 100 */
                c.hit = f3m(OR_NOHIT, OR_NOHIT, OR_NOHIT);
                c.clr[0] = c.clr[1] = c.clr[2] = c.clr[3] = 1;
                ray_cast(&c, pos, dir);
                if (c.hit.z != OR_NOHIT) {
                    o32[0] = c.hit.x * 0.01f; o32[1] = c.hit.y * 0.01f; o32[2] = c.hit.z * 0.01f; o32[3] = 1.0f;
                } else {
                    for (int k = 0; k < 4; k++) o32[k] = rp->backclr[k];
                }
                for (int k = 0; k < 4; k++) o64[k] = o32[k];
            } else if (mode == M_DEEP_LUT) {
                c.hit = f3m(0, 0, OR_NOHIT);
                c.clr[0] = c.clr[1] = c.clr[2] = 0; c.clr[3] = 1;
                c.clr64[0] = c.clr64[1] = c.clr64[2] = 0; c.clr64[3] = 1;
                ray_cast(&c, pos, dir);
                /* MOD.cu:75: (lerp3(back, clr, 1-T), 1-T) */
                float a = 1.0f - c.clr[3];
                double a64 = 1.0 - c.clr64[3];
                for (int k = 0; k < 3; k++) {
                    o32[k] = rp->backclr[k] + a * (c.clr[k] - rp->backclr[k]);
                    o64[k] = rp->backclr[k] + a64 * (c.clr64[k] - rp->backclr[k]);
                }
                o32[3] = a;
                o64[3] = a64;
            } else {
                c.hit = mode == M_SURF_LS ? f3m(0, 0, OR_NOHIT) : f3m(OR_NOHIT, OR_NOHIT, OR_NOHIT);
                c.clr[0] = c.clr[1] = c.clr[2] = c.clr[3] = 1;
                ray_cast(&c, pos, dir);
                phong(&c, pos, o32, o64);
            }
            int64_t px = (int64_t)y * cam->width + x;
            if (out32) for (int k = 0; k < 4; k++) out32[4 * px + k] = o32[k];
            if (out64) for (int k = 0; k < 4; k++) out64[4 * px + k] = o64[k];
        }
    return OR_OK;
}

/* Emission-absorption raycast with the linear transfer (extension): per
 * sample s, a = exp(extinct*s*step), I += s*T*(1-a)*albedo, T *= a. */
/* gvdbRaytrace MOD.cu:207-219 */
int or_raytrace(or_volume* v, int chan, const or_render_params* rp, or_scn_ray* rays, int64_t n, float bias) {
    or_camera one;
    memset(&one, 0, sizeof(one));
    one.width = one.height = 1;
    int st = render_check(v, chan, &one, rp);
    if (st) return st;
    if (n < 0 || (n > 0 && !rays)) return OR_ERR_ARG;
    for (int64_t i = 0; i < n; i++) {
        or_scn_ray* r = &rays[i];
        ray_ctx c;
        memset(&c, 0, sizeof(c));
        c.v = v; c.A = v->atlas[chan]; c.rp = rp; c.mode = M_SURF_TRI;
        c.hit = f3m(OR_NOHIT, OR_NOHIT, OR_NOHIT);
        c.norm = f3m(r->normal[0], r->normal[1], r->normal[2]);
        c.clr[0] = c.clr[1] = c.clr[2] = c.clr[3] = 1;
        f3 o = f3m(r->orig[0], r->orig[1], r->orig[2]), d = f3m(r->dir[0], r->dir[1], r->dir[2]);
        ray_cast(&c, o, d);
        if (c.hit.z != OR_NOHIT) c.hit = f3sub(c.hit, f3scale(d, bias));
        r->hit[0] = c.hit.x; r->hit[1] = c.hit.y; r->hit[2] = c.hit.z;
        r->normal[0] = c.norm.x; r->normal[1] = c.norm.y; r->normal[2] = c.norm.z;
    }
    return OR_OK;
}

int or_raycast_ea(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp, const double* bricks64,
                  float* I32, float* T32, double* I64, double* T64) {
    int st = render_check(v, chan, cam, rp);
    if (st) return st;
    for (int y = 0; y < cam->height; y++)
        for (int x = 0; x < cam->width; x++) {
            ray_ctx c;
            memset(&c, 0, sizeof(c));
            c.v = v; c.A = v->atlas[chan]; c.B64 = bricks64; c.rp = rp; c.mode = M_DEEP_LINEAR;
            c.hit = f3m(0, 0, OR_NOHIT);
            c.clr[3] = 1; c.clr64[3] = 1;
            f3 pos, dir;
            camera_ray(cam, x, y, &pos, &dir);
            ray_cast(&c, pos, dir);
            int64_t px = (int64_t)y * cam->width + x;
            if (I32) I32[px] = c.clr[0];
            if (T32) T32[px] = c.clr[3];
            if (I64) I64[px] = c.clr64[0];
            if (T64) T64[px] = c.clr64[3];
        }
    return OR_OK;
}

/* Adjoint of or_raycast_ea in FP64 by reverse accumulation over the recorded
 * samples of each ray: with e = extinct*step, a_k = exp(e s_k),
 *   dI/ds_k = albedo T_k ((1-a_k) - e s_k a_k) + e * sum_{j>k} albedo s_j T_j (1-a_j)
 *   dT/ds_k = e T_final
 * and ds_k/dvoxel = trilinear weight.  grad has nb*br^3 entries (padded
 * bricks, leaf-major, x fastest) and is accumulated into. */
/* mag != 0: the same sums with every term replaced by the magnitude of its
 * operands (|gI|, |gT|, |1 - a| + |e s a|, |e| times the magnitude of the
 * whole I sum instead of I_final - I_k); the result bounds the rounding
 * error of an FP32 evaluation of the adjoint, entry by entry, as (roundings
 * on the longest dependent chain) * 2^-24 * result, provided that
 * evaluation forms 1 - a to a few ulps of itself (from the series of
 * -expm1 near a = 1, not by subtracting a rounded a from 1). */
static int ea_backward_impl(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp,
                            const double* bricks64, const double* gI, const double* gT, double* grad, int mag) {
    int st = render_check(v, chan, cam, rp);
    if (st) return st;
    if (!grad) return OR_ERR_ARG;
    int br = brick_res(v);
    int64_t br3 = ipow3(br);
    double e = (double)rp->extinct * (double)rp->step;
    double alb = rp->albedo;
    ray_ctx c;
    memset(&c, 0, sizeof(c));
    c.caprec = 0;
    for (int y = 0; y < cam->height; y++)
        for (int x = 0; x < cam->width; x++) {
            c.v = v; c.A = v->atlas[chan]; c.B64 = bricks64; c.rp = rp; c.mode = M_DEEP_LINEAR;
            c.hit = f3m(0, 0, OR_NOHIT);
            c.clr[0] = c.clr[1] = c.clr[2] = 0; c.clr[3] = 1;
            c.clr64[0] = c.clr64[1] = c.clr64[2] = 0; c.clr64[3] = 1;
            c.nrec = 0;
            if (c.caprec == 0) { c.caprec = 256; c.rec = (or_sample*)malloc(256 * sizeof(or_sample)); }
            f3 pos, dir;
            camera_ray(cam, x, y, &pos, &dir);
            ray_cast(&c, pos, dir);
            int64_t px = (int64_t)y * cam->width + x;
            double gi = gI ? gI[px] : 0.0, gt = gT ? gT[px] : 0.0;
            int64_t n = c.nrec;
            double* Tk = (double*)malloc((size_t)(n + 1) * sizeof(double));
            double* sk = (double*)malloc((size_t)(n + 1) * sizeof(double));
            double T = 1.0;
            for (int64_t k = 0; k < n; k++) {
                sk[k] = tri64(&c, c.rec[k].leaf, c.rec[k].p);
                Tk[k] = T;
                T *= exp(e * sk[k]);
            }
            double Tfin = T, R = 0.0, Iabs = 0.0;
            if (mag)
                for (int64_t k = 0; k < n; k++) Iabs += fabs(alb * sk[k] * Tk[k] * (1.0 - exp(e * sk[k])));
            for (int64_t k = n - 1; k >= 0; k--) {
                double a = exp(e * sk[k]);
                double q = gi * (alb * Tk[k] * ((1.0 - a) - e * sk[k] * a) + e * R) + gt * e * Tfin;
                if (mag)
                    q = fabs(gi) * (fabs(alb) * Tk[k] * (fabs(1.0 - a) + fabs(e * sk[k] * a)) + fabs(e) * Iabs) +
                        fabs(gt * e) * Tfin;
                R += alb * sk[k] * Tk[k] * (1.0 - a);
                tri_cell tc = tri_locate(c.rec[k].p);
                double w[8];
                tri_weights64(&tc, w);
                for (int m = 0; m < 8; m++) {
                    int i = tc.i[0] + (m & 1), j = tc.i[1] + ((m >> 1) & 1), kk = tc.i[2] + (m >> 2);
                    i = i < 0 ? 0 : (i >= br ? br - 1 : i);
                    j = j < 0 ? 0 : (j >= br ? br - 1 : j);
                    kk = kk < 0 ? 0 : (kk >= br ? br - 1 : kk);
                    grad[c.rec[k].leaf * br3 + ((int64_t)kk * br + j) * br + i] += q * w[m];
                }
            }
            free(Tk);
            free(sk);
        }
    free(c.rec);
    return OR_OK;
}

int or_raycast_ea_backward(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp,
                           const double* bricks64, const double* gI, const double* gT, double* grad) {
    return ea_backward_impl(v, chan, cam, rp, bricks64, gI, gT, grad, 0);
}

int or_raycast_ea_backward_mag(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp,
                               const double* bricks64, const double* gI, const double* gT, double* grad) {
    return ea_backward_impl(v, chan, cam, rp, bricks64, gI, gT, grad, 1);
}

/* number of recorded samples per ray (for sample-set tests) */
int or_raycast_ea_samples(or_volume* v, int chan, const or_camera* cam, const or_render_params* rp, int64_t px,
                          int64_t max_out, int64_t* leaf_out, float* p_out, int64_t* count) {
    int st = render_check(v, chan, cam, rp);
    if (st) return st;
    if (px < 0 || px >= (int64_t)cam->width * cam->height) return OR_ERR_ARG;
    ray_ctx c;
    memset(&c, 0, sizeof(c));
    c.v = v; c.A = v->atlas[chan]; c.rp = rp; c.mode = M_DEEP_LINEAR;
    c.hit = f3m(0, 0, OR_NOHIT);
    c.clr[3] = 1; c.clr64[3] = 1;
    c.caprec = 256; c.rec = (or_sample*)malloc(256 * sizeof(or_sample));
    f3 pos, dir;
    camera_ray(cam, (int)(px % cam->width), (int)(px / cam->width), &pos, &dir);
    ray_cast(&c, pos, dir);
    for (int64_t k = 0; k < c.nrec && k < max_out; k++) {
        if (leaf_out) leaf_out[k] = c.rec[k].leaf;
        if (p_out) { p_out[3 * k] = c.rec[k].p.x; p_out[3 * k + 1] = c.rec[k].p.y; p_out[3 * k + 2] = c.rec[k].p.z; }
    }
    *count = c.nrec;
    free(c.rec);
    return OR_OK;
}

/* ------------------------------------------------------------------ */
/* FP64 padded-brick helpers for the adjoint chain                     */
/* ------------------------------------------------------------------ */

/* source leaf and padded index of apron voxel (i,j,k) of leaf b, by the
 * lookup of UpdateApron (OP.cuh:100-125); returns -1 for boundval */
static int64_t apron_source(const or_volume* v, int64_t b, int i, int j, int k, int64_t* src) {
    const or_node* nd = &v->nodes[0][b];
    int64_t w[3] = {(int64_t)nd->mPos[0] + i - 1, (int64_t)nd->mPos[1] + j - 1, (int64_t)nd->mPos[2] + k - 1};
    float wp[3] = {(float)w[0] + 0.5f, (float)w[1] + 0.5f, (float)w[2] + 0.5f};
    int64_t L = point_leaf(v, wp);
    if (L < 0) return -1;
    const or_node* ln = &v->nodes[0][L];
    int br = brick_res(v);
    int64_t li = w[0] - ln->mPos[0] + 1, lj = w[1] - ln->mPos[1] + 1, lk = w[2] - ln->mPos[2] + 1;
    *src = L * ipow3(br) + (lk * br + lj) * br + li;
    return L;
}

static int is_apron(int i, int j, int k, int br) {
    return i == 0 || j == 0 || k == 0 || i == br - 1 || j == br - 1 || k == br - 1;
}

/* UpdateApron on an FP64 padded brick array (interiors given) */
int or_apron_f64(or_volume* v, double* bricks, double boundval) {
    if (!v || !bricks || v->apron != 1) return OR_ERR_ARG;
    int br = brick_res(v);
    int64_t br3 = ipow3(br);
    for (int64_t b = 0; b < v->ncnt[0]; b++)
        for (int k = 0; k < br; k++)
            for (int j = 0; j < br; j++)
                for (int i = 0; i < br; i++) {
                    if (!is_apron(i, j, k, br)) continue;
                    int64_t src;
                    double val = apron_source(v, b, i, j, k, &src) < 0 ? boundval : bricks[src];
                    bricks[b * br3 + ((int64_t)k * br + j) * br + i] = val;
                }
    return OR_OK;
}

/* adjoint of or_apron_f64: each apron gradient is added to its source
 * interior voxel, then the apron gradients are cleared */
int or_apron_adjoint_f64(or_volume* v, double* grad) {
    if (!v || !grad || v->apron != 1) return OR_ERR_ARG;
    int br = brick_res(v);
    int64_t br3 = ipow3(br);
    for (int64_t b = 0; b < v->ncnt[0]; b++)
        for (int k = 0; k < br; k++)
            for (int j = 0; j < br; j++)
                for (int i = 0; i < br; i++) {
                    if (!is_apron(i, j, k, br)) continue;
                    int64_t src;
                    int64_t a = b * br3 + ((int64_t)k * br + j) * br + i;
                    if (apron_source(v, b, i, j, k, &src) >= 0) grad[src] += grad[a];
                }
    for (int64_t b = 0; b < v->ncnt[0]; b++)
        for (int k = 0; k < br; k++)
            for (int j = 0; j < br; j++)
                for (int i = 0; i < br; i++)
                    if (is_apron(i, j, k, br)) grad[b * br3 + ((int64_t)k * br + j) * br + i] = 0.0;
    return OR_OK;
}

/* adjoint of the weighted cubic splat in FP64: for the interior gradient g
 * (padded layout; apron entries ignored),
 *   dL/dw_i = sum_v g(v) t^3,  dL/dp_i = sum_v g(v) w_i 6 t^2 (v - p_i) / r^2,
 * t = 1 - |v-p_i|^2/r^2 over voxels of active leaves with |v-p_i| < r.
 * Positions are taken as doubles (pos64, 3 per point) so that finite
 * differences can perturb them below FP32 resolution. */
static int splat_backward_impl(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                               const double* grad, double* dpos, double* dw, int mag) {
    if (!v || (!pos64 && n > 0) || !grad) return OR_ERR_ARG;
    int br = brick_res(v);
    int64_t br3 = ipow3(br);
    double R = radius, R2 = R * R;
    for (int64_t i = 0; i < n; i++) {
        const double* w = pos64 + 3 * i;
        double wt = weights ? weights[i] : 1.0;
        if (mag) wt = fabs(wt);
        double gw = 0.0, gp[3] = {0, 0, 0};
        if (!(w[0] >= 0 && w[1] >= 0 && w[2] >= 0)) { if (dw) dw[i] = 0; if (dpos) dpos[3*i] = dpos[3*i+1] = dpos[3*i+2] = 0; continue; }
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)floor(w[a] - R);
            hi[a] = (int64_t)ceil(w[a] + R);
            if (lo[a] < 0) lo[a] = 0;
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    double dx = x - w[0], dy = y - w[1], dz = z - w[2];
                    double d2 = dx * dx + dy * dy + dz * dz;
                    if (!(d2 < R2)) continue;
                    int64_t leaf = voxel_leaf(v, x, y, z);
                    if (leaf < 0) continue;
                    const or_node* nd = &v->nodes[0][leaf];
                    int64_t lx = x - nd->mPos[0] + 1, ly = y - nd->mPos[1] + 1, lz = z - nd->mPos[2] + 1;
                    double g = grad[leaf * br3 + (lz * br + ly) * br + lx];
                    if (mag) { g = fabs(g); dx = fabs(dx); dy = fabs(dy); dz = fabs(dz); }
                    double t = 1.0 - d2 / R2;
                    gw += g * t * t * t;
                    double f = g * wt * 6.0 * t * t / R2;
                    gp[0] += f * dx; gp[1] += f * dy; gp[2] += f * dz;
                }
        if (dw) dw[i] = gw;
        if (dpos) { dpos[3 * i] = gp[0]; dpos[3 * i + 1] = gp[1]; dpos[3 * i + 2] = gp[2]; }
    }
    return OR_OK;
}

int or_splat_cubic_backward_f64(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                                const double* grad, double* dpos, double* dw) {
    return splat_backward_impl(v, pos64, n, weights, radius, grad, dpos, dw, 0);
}

/* the same sums over |g|, |w| and |v - p| per axis: a bound on the
 * magnitude of every term, for rounding-error floors */
int or_splat_cubic_backward_mag_f64(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                                    const double* grad, double* dpos, double* dw) {
    return splat_backward_impl(v, pos64, n, weights, radius, grad, dpos, dw, 1);
}

/* forward cubic splat in FP64 from FP64 positions into padded bricks
 * (interiors only; aprons left as they are) */
int or_splat_cubic_f64_padded(or_volume* v, const double* pos64, int64_t n, const double* weights, float radius,
                              double* bricks) {
    if (!v || (!pos64 && n > 0) || !bricks) return OR_ERR_ARG;
    int br = brick_res(v);
    int64_t br3 = ipow3(br);
    int r0 = v->res[0];
    for (int64_t b = 0; b < v->ncnt[0]; b++)
        for (int k = 1; k <= r0; k++)
            for (int j = 1; j <= r0; j++)
                for (int i = 1; i <= r0; i++) bricks[b * br3 + ((int64_t)k * br + j) * br + i] = 0.0;
    double R = radius, R2 = R * R;
    for (int64_t i = 0; i < n; i++) {
        const double* w = pos64 + 3 * i;
        if (!(w[0] >= 0 && w[1] >= 0 && w[2] >= 0)) continue;
        double wt = weights ? weights[i] : 1.0;
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)floor(w[a] - R);
            hi[a] = (int64_t)ceil(w[a] + R);
            if (lo[a] < 0) lo[a] = 0;
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    double dx = x - w[0], dy = y - w[1], dz = z - w[2];
                    double d2 = dx * dx + dy * dy + dz * dz;
                    if (!(d2 < R2)) continue;
                    int64_t leaf = voxel_leaf(v, x, y, z);
                    if (leaf < 0) continue;
                    const or_node* nd = &v->nodes[0][leaf];
                    int64_t lx = x - nd->mPos[0] + 1, ly = y - nd->mPos[1] + 1, lz = z - nd->mPos[2] + 1;
                    double t = 1.0 - d2 / R2;
                    bricks[leaf * br3 + (lz * br + ly) * br + lx] += wt * t * t * t;
                }
    }
    return OR_OK;
}
