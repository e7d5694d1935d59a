// render.hip -- HDDA raycasting over the sparse tree (Render) and the
// emission-absorption adjoint, for gfx950.
//
// One thread per ray, 8x8-pixel tiles per wave (neighbouring rays walk the
// same bricks), 16x16 pixels per 256-thread workgroup.  The traversal is the
// reference's hierarchical DDA (rayCast RC.cuh:543-611, HDDAState
// DDA.cuh:38-91): empty-space skipping through the child lists down to
// level 1, then a brick function marches the 8^3 brick at t = step*k.
// Every float expression keeps the reference's operand order with no
// contraction, so the set of samples (positions, bricks, transfer-function
// indices, surface hits) is a fixed function of the inputs:
//   dot(a,b) = a.x*b.x + a.y*b.y + a.z*b.z; normalize(v) = v * (1/sqrtf(dot))
//   trilinear: c = p + 0.5, i = floor(c), a = c - i, lerp u*(1-a) + w*a in
//   x, then y, then z, over the leaf's padded brick (voxel i centred at
//   i + 0.5, the texel convention of RC.cuh:104-108; indices clamped to the
//   padded brick).  Hardware tex3D's 8-bit weights are not reproduced.
// An HDDA axis with a zero direction component never steps (the reference's
// mask*tDel is 0*inf = NaN there and stalls the ray).
//
// The adjoint of the emission-absorption march records, per (ray, brick)
// segment, the entry t, transmittance T and intensity I (the transmittance
// checkpoint), sorts the segments by (leaf, ray) with the radix sort, and
// lets one wave per leaf replay its segments from the checkpoints against
// the brick staged in LDS, accumulating d loss / d voxel in LDS: no stored
// samples, no global float atomics, and a fixed accumulation order.
#include "gvdbx/gvdbx_c.h"
#include "gvdbx_kernels.h"

namespace {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 normalize(V3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// pixel ray (gvdbRayDeep MOD.cu:66-71 with getViewRay GEO.cuh:54-63)
__device__ __forceinline__ void camera_ray(const RenderCam& c, int x, int y, V3& pos, V3& dir) {
    const float u = ((float)x + 0.5f) / (float)c.width;
    const float w = ((float)y + 0.5f) / (float)c.height;
    pos = v3((u * c.org_u[0] + w * c.org_v[0]) + c.org[0], (u * c.org_u[1] + w * c.org_v[1]) + c.org[1],
             (u * c.org_u[2] + w * c.org_v[2]) + c.org[2]);
    V3 d = v3((u * c.dir_u[0] + w * c.dir_v[0]) + c.dir[0], (u * c.dir_u[1] + w * c.dir_v[1]) + c.dir[1],
              (u * c.dir_u[2] + w * c.dir_v[2]) + c.dir[2]);
    dir = normalize(d);
}

// rayBoxIntersect GEO.cuh:85-98
__device__ __forceinline__ V3 ray_box(V3 pos, V3 dir, const float* mn, const float* mx) {
    float h0 = (mn[0] - pos.x) / dir.x, h1 = (mx[0] - pos.x) / dir.x;
    float h2 = (mn[1] - pos.y) / dir.y, h3 = (mx[1] - pos.y) / dir.y;
    float h4 = (mn[2] - pos.z) / dir.z, h5 = (mx[2] - pos.z) / dir.z;
    float h6 = fmaxf(fmaxf(fminf(h0, h1), fminf(h2, h3)), fminf(h4, h5));
    float h7 = fminf(fminf(fmaxf(h0, h1), fmaxf(h2, h3)), fmaxf(h4, h5));
    h6 = (h6 < 0) ? 0.0f : h6;
    return v3(h6, h7, (h7 < h6 || h7 < 0) ? GVDBX_NOHIT : 0.0f);
}

// child of a node (getChild ND.cuh:115-124); reads outside the allocated
// lists are "no child"; returns the child's pool index or -1
__device__ __forceinline__ int64_t get_child(const RenderTree& tv, uint64_t listid, int64_t b) {
    if (listid == GVDBX_ID_UNDEFL) return -1;
    const int clev = (int)((listid >> 8) & 0xFF);
    if (clev <= 0 || clev >= tv.levs) return -1;
    const int64_t len = (int64_t)1 << (3 * tv.logdim[clev]);
    const int64_t flat = (int64_t)(listid >> 16) * len + b;
    if (flat < 0 || flat >= tv.ncnt[clev] * len) return -1;
    const uint64_t e = tv.clist[clev][flat];
    if (e == GVDBX_ID_UNDEF64) return -1;
    return (int64_t)(int32_t)(uint32_t)(e >> 16);
}

struct Hdda {
    V3 pos, dir;
    int ps[3];
    float tdel[3];
    V3 t;
    int p[3];
    float ts[3];
    int mask[3];

    __device__ __forceinline__ void set(V3 po, V3 di, V3 t0) {
        pos = po;
        dir = di;
        ps[0] = di.x > 0 ? 1 : -1;
        ps[1] = di.y > 0 ? 1 : -1;
        ps[2] = di.z > 0 ? 1 : -1;
        t = t0;
    }
    __device__ __forceinline__ void prepare(V3 vmin, float vdel) {
        const V3 q = sub(add(pos, scale(dir, t.x)), vmin);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float pf = comp(q, a) / vdel;
            tdel[a] = fabsf(vdel / comp(dir, a));
            const float fl = floorf(pf);
            const float br = ((fl - pf) + 0.5f) * (float)ps[a] + 0.5f;
            ts[a] = __builtin_isinf(tdel[a]) ? __builtin_inff() : br * tdel[a] + t.x;
            p[a] = (int)fl;
        }
    }
    // PrepareLeaf DDA.cuh:67-73: vdel 1 and no "+ t.x" in tSide (only the
    // order of the steps matters inside a brick)
    __device__ __forceinline__ void prepare_leaf(V3 vmin) {
        const V3 q = sub(add(pos, scale(dir, t.x)), vmin);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float pf = comp(q, a);
            tdel[a] = fabsf(1.0f / comp(dir, a));
            const float fl = floorf(pf);
            const float br = ((fl - pf) + 0.5f) * (float)ps[a] + 0.5f;
            ts[a] = __builtin_isinf(tdel[a]) ? __builtin_inff() : br * tdel[a];
            p[a] = (int)fl;
        }
    }
    __device__ __forceinline__ void next() {
        mask[0] = (ts[0] < ts[1]) & (ts[0] <= ts[2]);
        mask[1] = (ts[1] < ts[2]) & (ts[1] <= ts[0]);
        mask[2] = (ts[2] < ts[0]) & (ts[2] <= ts[1]);
        t.y = mask[0] ? ts[0] : (mask[1] ? ts[1] : ts[2]);
    }
    __device__ __forceinline__ void step() {
        t.x = t.y;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (mask[a]) ts[a] = ts[a] + tdel[a];
            p[a] += mask[a] * ps[a];
        }
    }
};

// trilinear sample of a padded brick (global memory or LDS)
template <int BR>
__device__ __forceinline__ float trilinear(const float* B, int br, V3 p) {
    const int n = BR > 0 ? BR : br;
    const float cx = p.x + 0.5f, cy = p.y + 0.5f, cz = p.z + 0.5f;
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    const float ax = cx - fx, ay = cy - fy, az = cz - fz;
    const float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const int x0 = min(max(ix, 0), n - 1), x1 = min(max(ix + 1, 0), n - 1);
    const int y0 = min(max(iy, 0), n - 1), y1 = min(max(iy + 1, 0), n - 1);
    const int z0 = min(max(iz, 0), n - 1), z1 = min(max(iz + 1, 0), n - 1);
    const float* r00 = B + (z0 * n + y0) * n;
    const float* r10 = B + (z0 * n + y1) * n;
    const float* r01 = B + (z1 * n + y0) * n;
    const float* r11 = B + (z1 * n + y1) * n;
    const float c00 = r00[x0] * bx + r00[x1] * ax;
    const float c10 = r10[x0] * bx + r10[x1] * ax;
    const float c01 = r01[x0] * bx + r01[x1] * ax;
    const float c11 = r11[x0] * bx + r11[x1] * ax;
    const float c0 = c00 * by + c10 * ay;
    const float c1 = c01 * by + c11 * ay;
    return c0 * bz + c1 * az;
}

// trilinear<BR> for a sample strictly inside the brick's interior range
// (0 <= p < res0 on every axis, as the brick marches guarantee): the cell
// floor(p + 0.5) and its +1 neighbour are then inside the padded brick, so
// the eight corners sit at one base index plus compile-time offsets and no
// clamps are needed.  Same arithmetic as trilinear<BR>, so the same bits.
// Returns the sample and its cell / fractions for the adjoint.
template <int BR>
__device__ __forceinline__ float trilinear_in(const float* B, V3 p, int& ix, int& iy, int& iz, float& ax, float& ay,
                                              float& az) {
    static_assert(BR > 0, "compile-time brick size");
    const float cx = p.x + 0.5f, cy = p.y + 0.5f, cz = p.z + 0.5f;
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    ax = cx - fx, ay = cy - fy, az = cz - fz;
    const float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
    ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const float* b = B + (__mul24(iz, BR * BR) + __mul24(iy, BR) + ix);   // indices < 2^23
    const float c00 = b[0] * bx + b[1] * ax;
    const float c10 = b[BR] * bx + b[BR + 1] * ax;
    const float c01 = b[BR * BR] * bx + b[BR * BR + 1] * ax;
    const float c11 = b[BR * BR + BR] * bx + b[BR * BR + BR + 1] * ax;
    const float c0 = c00 * by + c10 * ay;
    const float c1 = c01 * by + c11 * ay;
    return c0 * bz + c1 * az;
}

// transfer DDA.cuh:20-23 (float quotient, double clamp and scale, truncation)
__device__ __forceinline__ int transfer_index(const RenderScene& sc, float s) {
    const float q = (s - sc.thresh) / (sc.vmax - sc.vmin);
    const double qc = fmin(1.0, fmax(0.0, (double)q));
    return (int)(qc * 16300.0);
}

enum { MODE_DEEP_LUT = 0, MODE_DEEP_LINEAR = 1, MODE_SURF_TRI = 2, MODE_SURF_LS = 3, MODE_SURF_VOXEL = 4,
       MODE_SURF_CUBIC = 5, MODE_EMPTY = 6 };

struct RayState {
    float clr[4];
    V3 hit, norm;
    float hclr[4];   // the colour channel at a surface hit (Render's shading)
};

// getColorF RC.cuh:206-209 with make_float4(uchar4) cuda_math.cuh:221-224:
// the colour channel's uchar4 at voxel floor(p) of the leaf (brick-local p,
// interior from 0), each byte / 255.  The reference truncates p + the
// brick's atlas offset in FP32; floor(p) is the same voxel except where
// that sum rounds up to the next integer.
__device__ __forceinline__ void color_at(const RenderTree& tv, const RenderScene& sc, int64_t leaf, V3 p,
                                         float (&out)[4]) {
    const int br = tv.brick_res, ap = (br - tv.res[0]) >> 1;
    const int qx = min(max((int)floorf(p.x) + ap, 0), br - 1);
    const int qy = min(max((int)floorf(p.y) + ap, 0), br - 1);
    const int qz = min(max((int)floorf(p.z) + ap, 0), br - 1);
    const uchar4 c = sc.clr[leaf * (int64_t)(br * br * br) + (qz * br + qy) * br + qx];
    out[0] = (float)c.x / 255.0f;
    out[1] = (float)c.y / 255.0f;
    out[2] = (float)c.z / 255.0f;
    out[3] = (float)c.w / 255.0f;
}


// getGradient RC.cuh:132-142 (sign -1) / getGradientLevelSet RC.cuh:146-157 (+1)
template <int BR>
__device__ __forceinline__ V3 gradient(const float* B, int br, V3 p, bool ls) {
    float lo, hi;
    V3 g;
    lo = trilinear<BR>(B, br, v3(p.x - 0.5f, p.y, p.z));
    hi = trilinear<BR>(B, br, v3(p.x + 0.5f, p.y, p.z));
    g.x = ls ? hi - lo : lo - hi;
    lo = trilinear<BR>(B, br, v3(p.x, p.y - 0.5f, p.z));
    hi = trilinear<BR>(B, br, v3(p.x, p.y + 0.5f, p.z));
    g.y = ls ? hi - lo : lo - hi;
    lo = trilinear<BR>(B, br, v3(p.x, p.y, p.z - 0.5f));
    hi = trilinear<BR>(B, br, v3(p.x, p.y, p.z + 0.5f));
    g.z = ls ? hi - lo : lo - hi;
    return normalize(g);
}

// getTricubic RC.cuh:32-98: 27 filtered taps at the integer points
// floor(p) - 1 + {0, 1, 2} per axis (each the trilinear sample there),
// weights ta^2, 2 ta tb, tb^2 with tb = frac(p)/2 + 1/4, x then y then z.
// Brick-local coordinates; taps past the padded brick clamp to it (the
// reference reads on into the neighbouring atlas brick).
template <int BR>
__device__ __forceinline__ float tricubic(const float* B, int br, V3 p) {
    const float flx = floorf(p.x), fly = floorf(p.y), flz = floorf(p.z);
    float ta2[3], tab[3], tb2[3];
    const float fr[3] = {p.x - flx, p.y - fly, p.z - flz};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float tb = fr[k] * 0.5f + 0.25f;
        const float ta = 1.0f - tb;
        ta2[k] = ta * ta;
        tb2[k] = tb * tb;
        tab[k] = ta * tb * 2.0f;
    }
    const float qx = flx - 1.0f, qy = fly - 1.0f, qz = flz - 1.0f;
    float jkl[3];
#pragma unroll
    for (int zz = 0; zz < 3; zz++) {
        float row[3];
#pragma unroll
        for (int yy = 0; yy < 3; yy++) {
            const float t0 = trilinear<BR>(B, br, v3(qx, qy + (float)yy, qz + (float)zz));
            const float t1 = trilinear<BR>(B, br, v3(qx + 1.0f, qy + (float)yy, qz + (float)zz));
            const float t2 = trilinear<BR>(B, br, v3(qx + 2.0f, qy + (float)yy, qz + (float)zz));
            row[yy] = t0 * ta2[0] + t1 * tab[0] + t2 * tb2[0];
        }
        jkl[zz] = row[0] * ta2[1] + row[1] * tab[1] + row[2] * tb2[1];
    }
    return jkl[0] * ta2[2] + jkl[1] * tab[2] + jkl[2] * tb2[2];
}

// getGradientTricubic RC.cuh:161-171 (negated, half-voxel central differences)
template <int BR>
__device__ __forceinline__ V3 gradient_cubic(const float* B, int br, V3 p) {
    const float vs = 0.5f;
    V3 g;
    g.x = (tricubic<BR>(B, br, add(p, v3(-vs, 0, 0))) - tricubic<BR>(B, br, add(p, v3(vs, 0, 0)))) / (2 * vs);
    g.y = (tricubic<BR>(B, br, add(p, v3(0, -vs, 0))) - tricubic<BR>(B, br, add(p, v3(0, vs, 0)))) / (2 * vs);
    g.z = (tricubic<BR>(B, br, add(p, v3(0, 0, -vs))) - tricubic<BR>(B, br, add(p, v3(0, 0, vs)))) / (2 * vs);
    return normalize(g);
}

// Brick functions: rayDeepBrick RC.cuh:485-533 (DEEP_*),
// raySurfaceTrilinearBrick RC.cuh:281-300, rayLevelSetBrick RC.cuh:389-410
// (its rayLevelSet repeats the caller's test at the same point and returns
// p + vmin at once).
// node corner in the camera's (cloud-local) coordinates
__device__ __forceinline__ V3 node_pos(const GNode* nodes, int64_t id, const int* sh) {
    const int4 q = *(const int4*)(&nodes[id]);   // mLev.. | mPos[0..2]
    return v3((float)(q.y - sh[0]), (float)(q.z - sh[1]), (float)(q.w - sh[2]));
}

// Attenuation of one sample, e^(extinct * s * step), as 2^(k s) with
// k = extinct * step * log2(e), on the hardware exp2 (about 1 ulp; rounding
// k s adds |k s| 2^-24 relative).  The forward and the replay of the adjoint
// both use it, so the replay sees the forward's values.
__device__ __forceinline__ float ea_log2_atten(const RenderScene& sc) {
    return (sc.extinct * sc.step) * 1.44269504088896341f;
}
__device__ __forceinline__ float ea_atten(float k, float s) { return __builtin_amdgcn_exp2f(k * s); }

// 1 - a for a = ea_atten(k, s), to a few ulps.  Subtracting the rounded a
// from 1 keeps only log2(1 / (1 - a)) good bits when a is close to 1 (a
// sparse density: 1 - a ~ 1e-4 loses ten); for |x| < 1/2, x = k s ln 2,
// the series of -expm1(x) = -x (1 + x/2 + x^2/6 + ... + x^7/8!) is used
// instead (truncation below 2^-26 relative there).
__device__ __forceinline__ float ea_one_minus(float k, float s, float a) {
    const float x = (k * s) * 0.693147180559945309f;
    float p = __builtin_fmaf(x, 1.0f / 40320.0f, 1.0f / 5040.0f);
    p = __builtin_fmaf(x, p, 1.0f / 720.0f);
    p = __builtin_fmaf(x, p, 1.0f / 120.0f);
    p = __builtin_fmaf(x, p, 1.0f / 24.0f);
    p = __builtin_fmaf(x, p, 1.0f / 6.0f);
    p = __builtin_fmaf(x, p, 0.5f);
    p = __builtin_fmaf(x, p, 1.0f);
    return __builtin_fabsf(x) < 0.5f ? -x * p : 1.0f - a;
}

// 0 <= q < r0 on every axis, as 3-way min / max (positions are never NaN):
// the same predicate as the six compares, in fewer instructions.
__device__ __forceinline__ bool in_brick(const V3& q, float r0) {
    return __builtin_fminf(__builtin_fminf(q.x, q.y), q.z) >= 0.0f &&
           __builtin_fmaxf(__builtin_fmaxf(q.x, q.y), q.z) < r0;
}

template <int MODE, int BR>
__device__ __forceinline__ void brick_fn(const RenderTree& tv, const RenderScene& sc, int64_t leaf, V3 t, V3 pos,
                                         V3 dir, RayState& st, const int* sh) {
    const int br = BR > 0 ? BR : tv.brick_res;
    const float* B = tv.chan + leaf * (int64_t)(br * br * br);
    const V3 vmin = node_pos(tv.nodes[0], leaf, sh);
    const float r0 = (float)tv.res[0];
    if (MODE == MODE_DEEP_LUT || MODE == MODE_DEEP_LINEAR) {
        const float step = sc.step;
        const float ek = ea_log2_atten(sc);
        t.x = step * ceilf(t.x / step);
        V3 p = sub(add(pos, scale(dir, t.x)), vmin);
        const V3 wpt = scale(dir, step);
        const float dt = sqrtf(dot(wpt, wpt));
        if (st.hit.x == 0) st.hit.x = t.x;
        for (int iter = 0; st.clr[3] > sc.alphacut && iter < sc.max_iter && in_brick(p, r0); iter++) {
            float s;
            if constexpr (MODE == MODE_DEEP_LINEAR && BR > 0) {
                int ix, iy, iz;
                float ax, ay, az;
                s = trilinear_in<BR>(B, p, ix, iy, iz, ax, ay, az);
            } else {
                s = trilinear<BR>(B, br, p);
            }
            if (s >= sc.minval) {
                float4 val;
                if (MODE == MODE_DEEP_LUT)
                    val = sc.lut[transfer_index(sc, s)];
                else
                    val = make_float4(s, s, s, s);
                const float om = ea_one_minus(ek, val.w, ea_atten(ek, val.w));
                val.w = ea_atten(ek, val.w);
                if (MODE == MODE_DEEP_LUT && sc.clr) {
                    // RGB from the colour channel at the sample (RC.cuh:519)
                    float h[4];
                    color_at(tv, sc, leaf, p, h);
                    st.clr[0] += val.x * st.clr[3] * om * sc.albedo * h[0];
                    st.clr[1] += val.y * st.clr[3] * om * sc.albedo * h[1];
                    st.clr[2] += val.z * st.clr[3] * om * sc.albedo * h[2];
                } else {
                    st.clr[0] += val.x * st.clr[3] * om * sc.albedo;
                    st.clr[1] += val.y * st.clr[3] * om * sc.albedo;
                    st.clr[2] += val.z * st.clr[3] * om * sc.albedo;
                }
                st.clr[3] *= val.w;
            }
            p = add(p, wpt);
            t.x += dt;
        }
        st.hit.y = t.x;
        if (MODE == MODE_DEEP_LUT) {
            st.clr[0] = fminf(st.clr[0], 1.f);
            st.clr[1] = fminf(st.clr[1], 1.f);
            st.clr[2] = fminf(st.clr[2], 1.f);
            st.clr[3] = fmaxf(st.clr[3], 0.f);
        }
    } else if (MODE == MODE_EMPTY) {
        st.hit = add(pos, scale(dir, t.x));   // rayEmptySkipBrick RC.cuh:425-428
    } else if (MODE == MODE_SURF_VOXEL) {
        // raySurfaceVoxelBrick RC.cuh:227-263: the leaf's voxels in DDA
        // order; the first above the threshold (its exact value: the texel
        // centre) is hit as a unit cube; a missed cube repeats the voxel
        // without a step until the iteration limit, as the reference does
        const int ap = (br - tv.res[0]) >> 1;
        const int r0i = tv.res[0];
        V3 vm = vmin;
        Hdda d;
        d.set(pos, dir, t);
        d.prepare_leaf(vm);
        for (int iter = 0; iter < sc.max_iter && d.p[0] >= 0 && d.p[1] >= 0 && d.p[2] >= 0 && d.p[0] < r0i &&
                           d.p[1] < r0i && d.p[2] < r0i;
             iter++) {
            if (B[((d.p[2] + ap) * br + (d.p[1] + ap)) * br + (d.p[0] + ap)] > sc.thresh) {
                vm = add(vm, v3((float)d.p[0], (float)d.p[1], (float)d.p[2]));
                const float lo[3] = {vm.x, vm.y, vm.z}, hi[3] = {vm.x + 1.0f, vm.y + 1.0f, vm.z + 1.0f};
                const V3 tb = ray_box(pos, dir, lo, hi);
                d.t = tb;
                if (tb.z == GVDBX_NOHIT) {
                    st.hit.z = GVDBX_NOHIT;
                    continue;
                }
                const V3 hit = add(pos, scale(dir, tb.x));
                V3 fc = sub(sub(hit, vm), v3(0.5f, 0.5f, 0.5f));
                fc = sub(fc, scale(dir, 0.01f));
                const float m = fmaxf(fmaxf(fabsf(fc.x), fabsf(fc.y)), fabsf(fc.z));
                st.norm = v3(fabsf(fc.x) == m ? copysignf(1.0f, fc.x) : 0.0f, fabsf(fc.y) == m ? copysignf(1.0f, fc.y) : 0.0f,
                             fabsf(fc.z) == m ? copysignf(1.0f, fc.z) : 0.0f);
                st.hit = hit;
                if (sc.clr) color_at(tv, sc, leaf, v3((float)d.p[0], (float)d.p[1], (float)d.p[2]), st.hclr);
                return;
            }
            d.next();
            d.step();
        }
    } else if (MODE == MODE_SURF_CUBIC) {
        // raySurfaceTricubicBrick RC.cuh:316-340
        V3 p = sub(add(pos, scale(dir, t.x)), vmin);
        for (int iter = 0; iter < sc.max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0;
             iter++) {
            const float vz = tricubic<BR>(B, br, p);
            if (vz >= sc.thresh) {
                const float vx = tricubic<BR>(B, br, sub(p, scale(dir, sc.finestep)));
                const float vy = (vz - sc.thresh) / (vz - vx);
                p = add(p, scale(dir, -vy * sc.finestep));
                st.hit = add(p, vmin);
                st.norm = gradient_cubic<BR>(B, br, p);
                if (sc.clr) color_at(tv, sc, leaf, p, st.hclr);
                return;
            }
            p = add(p, scale(dir, sc.step));
        }
    } else if (MODE == MODE_SURF_TRI) {
        t.x = sc.step * ceilf(t.x / sc.step);
        V3 p = sub(add(pos, scale(dir, t.x)), vmin);
        for (int iter = 0; iter < sc.max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < r0 && p.y < r0 && p.z < r0;
             iter++) {
            if (trilinear<BR>(B, br, p) >= sc.thresh) {
                st.hit = add(p, vmin);
                st.norm = gradient<BR>(B, br, p, false);
                if (sc.clr) color_at(tv, sc, leaf, p, st.hclr);
                return;
            }
            p = add(p, scale(dir, sc.step));
        }
    } else {
        V3 p = sub(add(pos, scale(dir, t.x)), vmin);
        for (int iter = 0;
             iter < sc.max_iter && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x <= r0 && p.y <= r0 && p.z <= r0; iter++) {
            if (trilinear<BR>(B, br, p) < sc.thresh) {
                st.hit = add(p, vmin);
                st.norm = gradient<BR>(B, br, p, true);
                if (sc.clr) color_at(tv, sc, leaf, p, st.hclr);
                return;
            }
            p = add(p, scale(dir, sc.step));
        }
    }
}

// Segment checkpoints are buffered per wave in LDS and written out as chunks
// of up to VIS_BUF, so a wave takes one slot range (one global atomic) per
// chunk instead of one per march step.  Where a chunk lands depends on which
// wave flushes first, so each chunk is also described by a tag (global wave
// index, records the wave wrote before it), and every wave leaves its record
// total in wave_recs: the backward places chunk c at (exclusive scan of
// wave_recs)[wave] + tag's record offset, which lays the records out in
// (wave, march) order whichever wave wrote first, and so fixes the order of
// every sum over them.
constexpr int VIS_BUF = 448;   // ~9 KB per wave, 4 waves per block: 4 blocks fit a CU, as many as the registers allow
struct VisitLds {
    uint32_t key[VIS_BUF];
    float4 rec[VIS_BUF];
    int cnt;
    int pre;       // records written so far in earlier chunks
    uint32_t wave; // global wave index (chunk tags)
};

// Write the cnt buffered checkpoints as one chunk, with the lanes in `act`
// (any subset of the wave).
__device__ __forceinline__ void flush_visits(VisitLds& vl, const VisitRec& vr, int cnt, uint64_t act) {
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int n = __popcll(act);
    const int rank = __popcll(act & ((1ull << lane) - 1));
    uint32_t base = 0;
    if (lane == leader) {
        // one atomic takes the record range (low word) and the chunk slot (high word)
        const unsigned long long old =
            atomicAdd((unsigned long long*)vr.count, (unsigned long long)cnt | (1ull << 32));
        base = (uint32_t)old;
        const uint32_t cid = (uint32_t)(old >> 32);
        const int pre = __hip_atomic_load(&vl.pre, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        if ((int64_t)cid < vr.chunk_cap) {
            vr.chunk_tag[cid] = ((uint64_t)vl.wave << 32) | (uint32_t)pre;
            vr.chunk_base[cid] = base;
            vr.chunk_size[cid] = (uint32_t)cnt;
        }
        __hip_atomic_store(&vl.pre, pre + cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    base = __builtin_amdgcn_readlane(base, leader);
    __builtin_amdgcn_wave_barrier();
    for (int i = rank; i < cnt; i += n) {
        const int64_t slot = (int64_t)base + i;
        if (slot < vr.cap) {
            vr.key[slot] = vl.key[i];
            vr.rec[slot] = vl.rec[i];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// The buffer fill count is shared by the wave's lanes (read by all, written
// by one): wavefront-scope atomics keep every access in program order.
__device__ __forceinline__ int visit_count(VisitLds& vl) {
    return __hip_atomic_load(&vl.cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ void set_visit_count(VisitLds& vl, int c) {
    __hip_atomic_store(&vl.cnt, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ void record_visit(VisitLds& vl, const VisitRec& vr, int64_t leaf, uint32_t ray, float t,
                                             float T, float I) {
    const uint64_t act = __ballot(1);
    const int lane = __lane_id();
    const int n = __popcll(act);
    const int rank = __popcll(act & ((1ull << lane) - 1));
    int cnt = visit_count(vl);
    if (cnt + n > VIS_BUF) {
        flush_visits(vl, vr, cnt, act);
        cnt = 0;
    }
    vl.key[cnt + rank] = (uint32_t)leaf;
    vl.rec[cnt + rank] = make_float4(t, T, I, __uint_as_float(ray));
    __builtin_amdgcn_wave_barrier();
    if (rank == 0) set_visit_count(vl, cnt + n);
}

template <int MODE, int BR, bool RECORD>
__device__ void ray_cast(const RenderTree& tv, const RenderScene& sc, const RenderCam& cam, V3 pos, V3 dir,
                         RayState& st, uint32_t ray, const VisitRec& vr, VisitLds* vl = nullptr) {
    const int top = tv.top_lev;
    if (top < 1 || top >= tv.levs || tv.ncnt[top] < 1) return;
    const int* sh = cam.shift;
    int64_t nodeid[GVDBX_MAXLEV];
    float tmax[GVDBX_MAXLEV];
    int lev = top;
    nodeid[lev] = 0;
    V3 ts = cam.has_bounds ? ray_box(pos, dir, cam.bmin, cam.bmax) : ray_box(pos, dir, tv.bmin, tv.bmax);
    uint64_t list = tv.nodes[lev][0].mChildList;
    V3 vmin = node_pos(tv.nodes[lev], 0, sh);
    if (ts.z == GVDBX_NOHIT) return;
    ts.x += sc.epsilon;
    tmax[lev] = ts.y - sc.epsilon;
    Hdda d;
    d.set(pos, dir, ts);
    d.prepare(vmin, tv.vdel[lev]);
    for (int iter = 0; iter < sc.max_iter && lev > 0 && lev <= top && d.p[0] >= 0 && d.p[1] >= 0 && d.p[2] >= 0 &&
                       d.p[0] <= tv.res[lev] && d.p[1] <= tv.res[lev] && d.p[2] <= tv.res[lev];
         iter++) {
        d.next();
        const int dim = tv.logdim[lev];
        const int64_t b = ((((int64_t)d.p[2] << dim) + d.p[1]) << dim) + d.p[0];
        const int64_t child = get_child(tv, list, b);
        if (child >= 0 && child < tv.ncnt[lev - 1]) {
            if (lev == 1) {
                d.t.x += sc.epsilon;
                if (RECORD) record_visit(*vl, vr, child, ray, d.t.x, st.clr[3], st.clr[0]);
                brick_fn<MODE, BR>(tv, sc, child, d.t, pos, dir, st, sh);
                if (st.clr[3] <= 0) {
                    st.clr[3] = 0;
                    return;
                }
                if (st.hit.z != GVDBX_NOHIT) return;
                d.step();
            } else {
                lev--;
                nodeid[lev] = child;
                list = tv.nodes[lev][child].mChildList;
                vmin = node_pos(tv.nodes[lev], child, sh);
                d.t.x += sc.epsilon;
                tmax[lev] = d.t.y - sc.epsilon;
                d.prepare(vmin, tv.vdel[lev]);
            }
        } else {
            d.step();
        }
        while (lev <= top && d.t.x > tmax[lev]) {
            lev++;
            if (lev <= top) {
                list = tv.nodes[lev][nodeid[lev]].mChildList;
                vmin = node_pos(tv.nodes[lev], nodeid[lev], sh);
                d.prepare(vmin, tv.vdel[lev]);
            }
        }
    }
}

// 16x16 pixels per workgroup, 8x8 per wave
__device__ __forceinline__ bool pixel_of(const RenderCam& cam, int& x, int& y) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    x = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    y = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    return x < cam.width && y < cam.height;
}

__device__ __forceinline__ uchar4 to_u8(float4 c) {
    auto q = [](float v) -> unsigned char {
        float s = v * 255.0f;
        s = fminf(fmaxf(s, 0.0f), 255.0f);
        return (unsigned char)(int)s;
    };
    return make_uchar4(q(c.x), q(c.y), q(c.z), q(c.w));
}

// gvdbRayDeep MOD.cu:60-78
template <int BR>
__global__ __launch_bounds__(256) void k_render_volume(RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba,
                                                       uchar4* rgba8) {
    int x, y;
    if (!pixel_of(cam, x, y)) return;
    V3 pos, dir;
    camera_ray(cam, x, y, pos, dir);
    RayState st;
    st.clr[0] = st.clr[1] = st.clr[2] = 0.0f;
    st.clr[3] = 1.0f;
    st.hit = v3(0, 0, GVDBX_NOHIT);
    VisitRec none{};
    ray_cast<MODE_DEEP_LUT, BR, false>(tv, sc, cam, pos, dir, st, 0, none);
    const float a = 1.0f - st.clr[3];
    float4 o;
    o.x = sc.backclr[0] + a * (st.clr[0] - sc.backclr[0]);
    o.y = sc.backclr[1] + a * (st.clr[1] - sc.backclr[1]);
    o.z = sc.backclr[2] + a * (st.clr[2] - sc.backclr[2]);
    o.w = a;
    const int64_t px = (int64_t)y * cam.width + x;
    if (rgba) rgba[px] = o;
    if (rgba8) rgba8[px] = to_u8(o);
}

// gvdbRaySurfaceTrilinear / gvdbRayLevelSet MOD.cu:101-181 with
// performPhongShading MOD.cu:38-57: the hit's colour (1 without a colour
// channel) times diffuse + ambient
template <int MODE, int BR>
__global__ __launch_bounds__(256) void k_render_surface(RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba,
                                                        uchar4* rgba8) {
    int x, y;
    if (!pixel_of(cam, x, y)) return;
    V3 pos, dir;
    camera_ray(cam, x, y, pos, dir);
    RayState st;
    st.clr[0] = st.clr[1] = st.clr[2] = st.clr[3] = 1.0f;
    st.hit = MODE == MODE_SURF_LS ? v3(0, 0, GVDBX_NOHIT) : v3(GVDBX_NOHIT, GVDBX_NOHIT, GVDBX_NOHIT);
    st.norm = v3(0, 0, 0);
    st.hclr[0] = st.hclr[1] = st.hclr[2] = st.hclr[3] = 1.0f;
    VisitRec none{};
    ray_cast<MODE, BR, false>(tv, sc, cam, pos, dir, st, 0, none);
    float4 o;
    if (st.hit.z == GVDBX_NOHIT) {
        o = make_float4(sc.backclr[0], sc.backclr[1], sc.backclr[2], sc.backclr[3]);
    } else {
        const V3 light = v3(sc.light[0], sc.light[1], sc.light[2]);
        const V3 ldir = normalize(sub(light, st.hit));
        float diff = (float)(0.9 * (double)fmaxf(0.0f, dot(st.norm, ldir)));
        const float amb = 0.1f;
        if (sc.shadow_amt > 0) {
            RayState s2;
            s2.clr[0] = s2.clr[1] = s2.clr[2] = 0.0f;
            s2.clr[3] = 1.0f;
            s2.hit = v3(0, 0, GVDBX_NOHIT);
            s2.norm = v3(0, 0, 0);
            // gvdbRaySurfaceTricubic shades with the trilinear brick (MOD.cu:137)
            constexpr int SMODE = MODE == MODE_SURF_CUBIC ? MODE_SURF_TRI : MODE;
            ray_cast<SMODE, BR, false>(tv, sc, cam, add(st.hit, scale(st.norm, sc.shadow_bias)), ldir, s2, 0, none);
            if (s2.hit.z != GVDBX_NOHIT) diff = (float)((double)diff * (1.0 - (double)sc.shadow_amt));
        }
        const float sh = diff + amb;
        o = make_float4(st.hclr[0] * sh, st.hclr[1] * sh, st.hclr[2] * sh, 1.0f);
    }
    const int64_t px = (int64_t)y * cam.width + x;
    if (rgba) rgba[px] = o;
    if (rgba8) rgba8[px] = to_u8(o);
}

// getNodeAtPoint ND.cuh:199-226,243-253: float descent from the top node to
// the leaf holding w, -1 when a list on the way has no entry or the leaf
// found does not contain w
__device__ __forceinline__ int64_t point_leaf(const RenderTree& tv, const float w[3]) {
    int lev = tv.top_lev;
    if (lev < 1 || lev >= tv.levs || tv.ncnt[lev] < 1) return -1;
    const int zero[3] = {0, 0, 0};
    V3 vmin = node_pos(tv.nodes[lev], 0, zero);
    uint64_t list = tv.nodes[lev][0].mChildList;
    int64_t id = 0;
    while (lev > 0) {
        const float rg = tv.vdel[lev] * (float)tv.res[lev];
        if (w[0] < vmin.x || w[1] < vmin.y || w[2] < vmin.z || w[0] >= vmin.x + rg || w[1] >= vmin.y + rg ||
            w[2] >= vmin.z + rg)
            return -1;
        const float vdel = tv.vdel[lev];
        const int p0 = (int)((w[0] - vmin.x) / vdel), p1 = (int)((w[1] - vmin.y) / vdel), p2 = (int)((w[2] - vmin.z) / vdel);
        const int dim = tv.logdim[lev];
        const int64_t b = ((((int64_t)p2 << dim) + p1) << dim) + p0;
        lev--;
        const int64_t c = get_child(tv, list, b);
        if (c < 0 || c >= tv.ncnt[lev]) return -1;
        id = c;
        vmin = node_pos(tv.nodes[lev], c, zero);
        list = tv.nodes[lev][c].mChildList;
    }
    const float r0 = (float)tv.res[0];
    if (!(w[0] >= vmin.x && w[0] < vmin.x + r0 && w[1] >= vmin.y && w[1] < vmin.y + r0 && w[2] >= vmin.z &&
          w[2] < vmin.z + r0))
        return -1;
    return id;
}

// transfer(gvdb, v): the colour of the transfer function at the sample
__device__ __forceinline__ float4 transfer_lut(const RenderScene& sc, float s) { return sc.lut[transfer_index(sc, s)]; }

// uchar4 of the colour with alpha 255 (the empty-skip and section kernels)
__device__ __forceinline__ uchar4 to_u8_opaque(float4 c) {
    uchar4 q = to_u8(c);
    q.w = 255;
    return q;
}

// gvdbRayEmptySkip MOD.cu:184-204: the first brick entry point / 100
template <int BR>
__global__ __launch_bounds__(256) void k_render_empty(RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba,
                                                      uchar4* rgba8) {
    int x, y;
    if (!pixel_of(cam, x, y)) return;
    V3 pos, dir;
    camera_ray(cam, x, y, pos, dir);
    RayState st;
    st.clr[0] = st.clr[1] = st.clr[2] = st.clr[3] = 1.0f;
    st.hit = v3(GVDBX_NOHIT, GVDBX_NOHIT, GVDBX_NOHIT);
    st.norm = v3(0, 0, 0);
    VisitRec none{};
    ray_cast<MODE_EMPTY, BR, false>(tv, sc, cam, pos, dir, st, 0, none);
    const float4 o = st.hit.z != GVDBX_NOHIT ? make_float4(st.hit.x * 0.01f, st.hit.y * 0.01f, st.hit.z * 0.01f, 1.0f)
                                            : make_float4(sc.backclr[0], sc.backclr[1], sc.backclr[2], sc.backclr[3]);
    const int64_t px = (int64_t)y * cam.width + x;
    if (rgba) rgba[px] = o;
    if (rgba8) rgba8[px] = to_u8_opaque(o);
}

// gvdbSection2D MOD.cu:268-293: pixel (x, y) samples the slice point
// pnt + ((2x/w - 1, 0, 2y/h - 1) * norm), the offsets formed in double
template <int BR>
__global__ __launch_bounds__(256) void k_render_section2d(RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba,
                                                          uchar4* rgba8) {
    int x, y;
    if (!pixel_of(cam, x, y)) return;
    const float sx = (float)((double)(float)x * 2.0 / (double)cam.width - 1.0);
    const float sz = (float)((double)(float)y * 2.0 / (double)cam.height - 1.0);
    const float w[3] = {sc.slice_pnt[0] + sx * sc.slice_norm[0], sc.slice_pnt[1] + 0.0f * sc.slice_norm[1],
                        sc.slice_pnt[2] + sz * sc.slice_norm[2]};
    const int64_t leaf = point_leaf(tv, w);
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (leaf >= 0) {
        const int br = BR > 0 ? BR : tv.brick_res;
        const int zero[3] = {0, 0, 0};
        const V3 vmin = node_pos(tv.nodes[0], leaf, zero);
        const float s = trilinear<BR>(tv.chan + leaf * (int64_t)(br * br * br), br, v3(w[0] - vmin.x, w[1] - vmin.y, w[2] - vmin.z));
        const float4 c = transfer_lut(sc, s);
        o = make_float4(0.0f + c.w * (c.x - 0.0f), 0.0f + c.w * (c.y - 0.0f), 0.0f + c.w * (c.z - 0.0f), 1.0f);
    }
    const int64_t px = (int64_t)y * cam.width + x;
    if (rgba) rgba[px] = o;
    if (rgba8) rgba8[px] = to_u8_opaque(o);
}

// gvdbSection3D MOD.cu:222-265: the transfer colour where the ray meets the
// slice plane, blended over the trilinear surface found beyond it
template <int BR>
__global__ __launch_bounds__(256) void k_render_section3d(RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba,
                                                          uchar4* rgba8) {
    int x, y;
    if (!pixel_of(cam, x, y)) return;
    V3 pos, dir;
    camera_ray(cam, x, y, pos, dir);
    float clr[4] = {1.0f, 1.0f, 1.0f, 0.0f};
    const float* n = sc.slice_norm;
    const float* q = sc.slice_pnt;
    // rayPlaneIntersect GEO.cuh:67-71
    float t = ((q[0] - pos.x) * n[0] + (q[1] - pos.y) * n[1] + (q[2] - pos.z) * n[2]) /
              (dir.x * n[0] + dir.y * n[1] + dir.z * n[2]);
    t = t > 0 ? t : GVDBX_NOHIT;
    if (t > 0) {
        pos = add(pos, scale(dir, t));
        const float w[3] = {pos.x, pos.y, pos.z};
        const int64_t leaf = point_leaf(tv, w);
        if (leaf >= 0) {
            const int br = BR > 0 ? BR : tv.brick_res;
            const int zero[3] = {0, 0, 0};
            const V3 vmin = node_pos(tv.nodes[0], leaf, zero);
            const V3 pl = sub(pos, vmin);
            t = trilinear<BR>(tv.chan + leaf * (int64_t)(br * br * br), br, pl);
            const float4 c = transfer_lut(sc, t);
            clr[0] = c.x; clr[1] = c.y; clr[2] = c.z; clr[3] = c.w;
            if (sc.clr) {   // MOD.cu:249-252: times the colour there (alpha times 1)
                float h[4];
                color_at(tv, sc, leaf, pl, h);
                clr[0] *= h[0]; clr[1] *= h[1]; clr[2] *= h[2];
            }
        } else {
            t = 0;
        }
    }
    RayState st;
    st.clr[0] = st.clr[1] = st.clr[2] = st.clr[3] = 1.0f;
    st.hit = v3(GVDBX_NOHIT, GVDBX_NOHIT, GVDBX_NOHIT);
    st.norm = v3(0, 0, 0);
    st.hclr[0] = st.hclr[1] = st.hclr[2] = st.hclr[3] = 1.0f;
    VisitRec none{};
    ray_cast<MODE_SURF_TRI, BR, false>(tv, sc, cam, pos, dir, st, 0, none);
    float o[4];
    if (st.hit.z != GVDBX_NOHIT) {
        const V3 ldir = normalize(sub(v3(sc.light[0], sc.light[1], sc.light[2]), st.hit));
        const float ds = (t > sc.thresh) ? 1.0f : (float)(0.8 * (double)fmaxf(0.0f, dot(st.norm, ldir)));
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float a = st.hclr[k] * ds;   // the surface colour (1 without a colour channel)
            o[k] = a + clr[3] * (clr[k] - a);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = sc.backclr[k] + clr[3] * (clr[k] - sc.backclr[k]);
    }
    const float4 of = make_float4(o[0], o[1], o[2], o[3]);
    const int64_t px = (int64_t)y * cam.width + x;
    if (rgba) rgba[px] = of;
    if (rgba8) rgba8[px] = to_u8_opaque(of);
}

// Render(SHADE_OFF) VG.cpp:4340-4345: the image is cleared
__global__ __launch_bounds__(256) void k_render_off(int64_t n, float4* rgba, uchar4* rgba8) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (rgba) rgba[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (rgba8) rgba8[i] = make_uchar4(0, 0, 0, 0);
    }
}

// gvdbRaytrace MOD.cu:207-219: each ray's trilinear surface hit, pulled back
// by bias along the ray; the normal is written only on a hit
template <int BR>
__global__ __launch_bounds__(256) void k_raytrace(RenderTree tv, RenderScene sc, GScnRay* rays, int64_t n, float bias) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GScnRay& r = rays[i];
    const V3 pos = v3(r.orig[0], r.orig[1], r.orig[2]), dir = v3(r.dir[0], r.dir[1], r.dir[2]);
    RenderCam cam{};
    RayState st;
    st.clr[0] = st.clr[1] = st.clr[2] = st.clr[3] = 1.0f;
    st.hit = v3(GVDBX_NOHIT, GVDBX_NOHIT, GVDBX_NOHIT);
    st.norm = v3(r.normal[0], r.normal[1], r.normal[2]);
    VisitRec none{};
    ray_cast<MODE_SURF_TRI, BR, false>(tv, sc, cam, pos, dir, st, 0, none);
    if (st.hit.z != GVDBX_NOHIT) st.hit = sub(st.hit, scale(dir, bias));
    r.hit[0] = st.hit.x; r.hit[1] = st.hit.y; r.hit[2] = st.hit.z;
    r.normal[0] = st.norm.x; r.normal[1] = st.norm.y; r.normal[2] = st.norm.z;
}

template <int BR, bool RECORD>
__global__ __launch_bounds__(256) void k_raycast_ea(RenderTree tv, RenderCam cam0, const RenderCam* cams,
                                                    RenderScene sc, float* I, float* T, VisitRec vr) {
    __shared__ VisitLds vbuf[RECORD ? 4 : 1];
    VisitLds& vl = vbuf[RECORD ? threadIdx.x / 64 : 0];
    if (RECORD && __lane_id() == 0) {
        set_visit_count(vl, 0);
        __hip_atomic_store(&vl.pre, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        vl.wave = (uint32_t)((((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * 4 +
                             threadIdx.x / 64);
    }
    const RenderCam& cam = cams ? cams[blockIdx.z] : cam0;
    int x, y;
    const bool ok = pixel_of(cam0, x, y);
    if (!ok) return;
    V3 pos, dir;
    camera_ray(cam, x, y, pos, dir);
    RayState st;
    st.clr[0] = st.clr[1] = st.clr[2] = 0.0f;
    st.clr[3] = 1.0f;
    st.hit = v3(0, 0, GVDBX_NOHIT);
    const uint32_t ray = (uint32_t)((int64_t)blockIdx.z * cam0.width * cam0.height + (int64_t)y * cam0.width + x);
    ray_cast<MODE_DEEP_LINEAR, BR, RECORD>(tv, sc, cam, pos, dir, st, ray, vr, &vl);
    I[ray] = st.clr[0];
    T[ray] = st.clr[3];
    if (RECORD) {
        const int cnt = visit_count(vl);
        const uint64_t act = __ballot(1);
        if (cnt > 0) flush_visits(vl, vr, cnt, act);
        // waves whose lanes all left early keep the zero the host filled in
        if (__lane_id() == __ffsll((unsigned long long)act) - 1)
            vr.wave_recs[vl.wave] = (uint32_t)__hip_atomic_load(&vl.pre, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
}

// [begin, end) of each leaf's run in the sorted segment keys
__global__ void k_segments(const uint32_t* key, int64_t n, int32_t* seg, int64_t nleaf) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t l = key[i];
        if (l >= (uint64_t)nleaf) continue;
        if (i == 0 || key[i - 1] != l) seg[2 * l] = (int32_t)i;
        if (i == n - 1 || key[i + 1] != l) seg[2 * l + 1] = (int32_t)(i + 1);
    }
}

// One wave per chunk: its records go to consecutive places from its wave's
// start (wave_off, the exclusive scan of the waves' record totals) plus the
// records its wave wrote before it (the tag's low word).
__global__ __launch_bounds__(64) void k_chunk_order(const uint64_t* tag, int64_t nchunk, const uint32_t* chunk_base,
                                                    const uint32_t* chunk_size, const int32_t* wave_off,
                                                    const uint32_t* key, uint32_t* key_out, uint32_t* val_out) {
    const int64_t c = blockIdx.x;
    if (c >= nchunk) return;
    const uint64_t t = tag[c];
    const uint32_t base = chunk_base[c], size = chunk_size[c];
    const int64_t dst = (int64_t)(uint32_t)wave_off[t >> 32] + (uint32_t)t;
    for (uint32_t j = threadIdx.x; j < size; j += 64) {
        key_out[dst + j] = key[base + j];
        val_out[dst + j] = base + j;
    }
}

// Move the four corner sums a cell shares with its neighbour along axis A:
// up (the ray moved +1) takes the d=1 layer to d=0, down takes d=0 to d=1.
template <int A>
__device__ __forceinline__ void shift_corners(float* acc, bool up) {
#pragma unroll
    for (int m = 0; m < 8; m++) {
        if (m & (1 << A)) continue;
        const float lo = acc[m], hi = acc[m | (1 << A)];
        acc[m] = up ? hi : 0.0f;
        acc[m | (1 << A)] = up ? 0.0f : lo;
    }
}

// One workgroup per leaf: replay every segment of the leaf from its
// checkpoint and accumulate dL/dvoxel = sum_k q_k * trilinear weight, with
//   e = extinct*step, a_k = exp(extinct*s_k*step),
//   q_k = gI * (albedo T_k ((1-a_k) - e s_k a_k) + e (I_final - I_k)) + gT e T_final
// (I_k: intensity after sample k).  The leaf's segments (sorted by ray) go
// to the four waves by the parity class of their pixel, one segment per lane
// at a time.  Each wave adds into its own LDS copy; the four copies are
// summed in a fixed order at the end.
// A lane keeps the eight corner sums of its current trilinear cell in
// registers and flushes a corner when the ray leaves it.  A flush first
// claims the voxel in its wave's own claim array: the lane writes its lane
// number to claim[wave][voxel] and reads it back.  Only lanes of the same
// wave write that array, and all of them in the same instruction, so which
// lane reads its own number back (the one the LDS kept for that byte) and
// therefore which lane adds with a plain read-add-write and which add
// afterwards with an LDS atomic depends on that wave's lanes alone: the
// summation order inside each wave's gradient copy, and so every bit of the
// result, does not depend on how the waves are timed against each other.
// LDS float atomics are slow (tens of cycles per instruction), so the claim
// pays for itself whenever most flushes are alone on their voxel, which
// rays that march side by side make the common case.
template <int BR>
__global__ __launch_bounds__(256) void k_ea_backward(EABackParams p) {
    constexpr int NV = BR * BR * BR;
    constexpr int NW = 4;
    // One LDS block laid out by hand with the brick first: its base address is
    // then 0, so a trilinear cell's four corner pairs are one address plus
    // small immediate offsets.
    struct EALds {
        float dens[NV];
        float grad[NW][NV];
        int32_t wlist[NW][128];   // per wave: its segments waiting for a lane
        uint8_t claim[NW][NV];    // per wave: lane number of the last claim of each voxel
    };
    __shared__ EALds lds;
    float* const dens = lds.dens;
    float (*const grad)[NV] = lds.grad;
    uint8_t* const claim = lds.claim[threadIdx.x >> 6];
    int32_t (*const wlist)[128] = lds.wlist;
    const int64_t leaf = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const float* src = p.tv.chan + leaf * (int64_t)NV;
    for (int i = tid; i < NV; i += 256) {
        dens[i] = src[i];
#pragma unroll
        for (int w = 0; w < NW; w++) {
            grad[w][i] = 0.0f;
            lds.claim[w][i] = 0xFF;
        }
    }
    __syncthreads();
    const int32_t b0 = p.seg[2 * leaf], b1 = p.seg[2 * leaf + 1];
    float* g = grad[wave];
    const RenderScene& sc = p.sc;
    const float r0 = (float)p.tv.res[0];
    const float e = sc.extinct * sc.step;
    const float ek = ea_log2_atten(sc);
    const uint32_t npix = (uint32_t)p.cam.width * (uint32_t)p.cam.height;
    auto replay = [&](int32_t c) {
        const uint32_t id = p.idx[c];
        const float4 cp = p.rec[id];
        const uint32_t ray = __float_as_uint(cp.w);
        const RenderCam& cam = p.cams ? p.cams[ray / npix] : p.cam;
        const uint32_t pix = ray % npix;
        const V3 vmin = node_pos(p.tv.nodes[0], leaf, cam.shift);
        V3 pos, dir;
        camera_ray(cam, (int)(pix % (uint32_t)p.cam.width), (int)(pix / (uint32_t)p.cam.width), pos, dir);
        float T = cp.y, I = cp.z;
        const float Ifin = p.I[ray], Tfin = p.T[ray], gi = p.gI[ray];
        const float gt = p.gT ? p.gT[ray] : 0.0f;
        const float step = sc.step;
        const float tt = step * ceilf(cp.x / step);
        V3 q = sub(add(pos, scale(dir, tt)), vmin);
        const V3 wpt = scale(dir, step);
        // corner sums of the current trilinear cell, kept in registers while
        // the ray stays in it; a step of one cell along one axis keeps the
        // four shared corners, anything else flushes all eight
        float acc[8];
#pragma unroll
        for (int m = 0; m < 8; m++) acc[m] = 0.0f;
        int cx = 0, cy = 0, cz = 0;
        int cb = 0;   // brick index of the current cell's corner 0
        bool have = false;
        auto flush = [&](int mask_keep, int axis_bit) {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                if (((m >> axis_bit) & 1) == mask_keep) continue;
                if (acc[m] != 0.0f) {
                    // the cell is inside the padded brick (samples stay in the interior range);
                    // corner m is corner 0 plus a constant (an immediate offset of the LDS access)
                    const int vi = cb + ((m >> 2) * BR * BR + ((m >> 1) & 1) * BR + (m & 1));
                    // relaxed atomics: plain LDS byte stores and loads the compiler may not fold
                    __hip_atomic_store(&claim[vi], (uint8_t)lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    if (__hip_atomic_load(&claim[vi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) == (uint8_t)lane)
                        g[vi] = g[vi] + acc[m];
                    else
                        atomicAdd(&g[vi], acc[m]);
                }
                acc[m] = 0.0f;
            }
        };
        for (int iter = 0; T > sc.alphacut && iter < sc.max_iter && in_brick(q, r0); iter++) {
            // trilinear sample (the forward's value) sharing its cell and weights with the adjoint
            int ix, iy, iz;
            float ax, ay, az;
            const float s = trilinear_in<BR>(dens, q, ix, iy, iz, ax, ay, az);
            const float bx = 1.0f - ax, by = 1.0f - ay, bz = 1.0f - az;
            if (s >= sc.minval) {
                const float a = ea_atten(ek, s);
                const float om = ea_one_minus(ek, s, a);
                I += s * T * om * sc.albedo;
                const float gq = gi * (sc.albedo * T * (om - e * s * a) + e * (Ifin - I)) + gt * e * Tfin;
                T *= a;
                if (have && !(ix == cx && iy == cy && iz == cz)) {
                    const int ddx = ix - cx, ddy = iy - cy, ddz = iz - cz;
                    int axis = -1, sgn = 0;
                    if (ddy == 0 && ddz == 0 && (ddx == 1 || ddx == -1)) axis = 0, sgn = ddx;
                    else if (ddx == 0 && ddz == 0 && (ddy == 1 || ddy == -1)) axis = 1, sgn = ddy;
                    else if (ddx == 0 && ddy == 0 && (ddz == 1 || ddz == -1)) axis = 2, sgn = ddz;
                    if (axis < 0) {
                        flush(-1, 0);   // keep nothing
                    } else {
                        // +1: old corners at d=1 become the new cell's d=0 (flush d=0, shift down)
                        // -1: old corners at d=0 become the new cell's d=1 (flush d=1, shift up)
                        const int keep = sgn > 0 ? 1 : 0;
                        flush(keep, axis);
                        if (axis == 0) shift_corners<0>(acc, keep != 0);
                        else if (axis == 1) shift_corners<1>(acc, keep != 0);
                        else shift_corners<2>(acc, keep != 0);
                    }
                }
                if (!have || !(ix == cx && iy == cy && iz == cz)) cb = __mul24(iz, BR * BR) + __mul24(iy, BR) + ix;
                cx = ix, cy = iy, cz = iz;
                have = true;
                // gradient only (no decision depends on it): fused products
                const float gz0 = gq * bz, gz1 = gq * az;
                const float gyz[4] = {gz0 * by, gz0 * ay, gz1 * by, gz1 * ay};
#pragma unroll
                for (int m = 0; m < 8; m++) acc[m] = __builtin_fmaf(gyz[m >> 1], (m & 1) ? ax : bx, acc[m]);
            }
            q = add(q, wpt);
        }
        if (have) flush(-1, 0);
    };
    // Wave w replays the leaf's segments whose pixel has parity class
    // (x & 1) | (y & 1) << 1 == w, in segment order, 64 at a time: rays of
    // one wave are then at least two pixels apart and rarely flush the same
    // voxel in one instruction.
    int cnt = 0;
    int32_t* wl = wlist[wave];
    for (int32_t j0 = b0; j0 < b1; j0 += 64) {
        const int32_t c = j0 + lane;
        bool mine = false;
        if (c < b1) {
            const uint32_t pix = __float_as_uint(p.rec[p.idx[c]].w) % npix;
            mine = (int)(((pix % (uint32_t)p.cam.width) & 1u) | (((pix / (uint32_t)p.cam.width) & 1u) << 1)) == wave;
        }
        const uint64_t mb = __ballot(mine);
        if (mine) wl[cnt + (int)__popcll(mb & ((1ull << lane) - 1ull))] = c;
        cnt += (int)__popcll(mb);
        if (cnt >= 64) {
            __builtin_amdgcn_wave_barrier();
            replay(wl[lane]);
            const int32_t rest = lane < cnt - 64 ? wl[64 + lane] : 0;
            __builtin_amdgcn_wave_barrier();
            if (lane < cnt - 64) wl[lane] = rest;
            cnt -= 64;
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < cnt) replay(wl[lane]);
    __syncthreads();
    float* dst = p.grad + leaf * (int64_t)NV;
    for (int i = tid; i < NV; i += 256) dst[i] = ((grad[0][i] + grad[1][i]) + grad[2][i]) + grad[3][i];
}

// Adjoint of UpdateApron: interior voxel (x,y,z) of leaf k is copied into the
// apron of the neighbour in direction (dx,dy,dz) at (x,y,z) - r0*(dx,dy,dz)
// whenever that lands in the neighbour's apron; gather those gradients in a
// fixed neighbour order.  Reads only apron entries and writes only interior
// entries, so it runs in place.
template <int BR>
__global__ __launch_bounds__(256) void k_apron_adjoint(ApronParams ap, float* grad) {
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
    float* brick = grad + k * (int64_t)nv;
    for (int i = threadIdx.x; i < r0 * r0 * r0; i += blockDim.x) {
        const int x = a + i % r0, y = a + (i / r0) % r0, z = a + i / (r0 * r0);
        const int lx = x == a ? -1 : 0, hx = x == a + r0 - 1 ? 1 : 0;
        const int ly = y == a ? -1 : 0, hy = y == a + r0 - 1 ? 1 : 0;
        const int lz = z == a ? -1 : 0, hz = z == a + r0 - 1 ? 1 : 0;
        if ((lx | hx | ly | hy | lz | hz) == 0) continue;
        float acc = 0.0f;
        for (int dz = lz; dz <= hz; dz++)
            for (int dy = ly; dy <= hy; dy++)
                for (int dx = lx; dx <= hx; dx++) {
                    if ((dx | dy | dz) == 0) continue;
                    const int32_t n = nbl[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
                    if (n < 0) continue;
                    acc += grad[(int64_t)n * nv + ((int64_t)(z - dz * r0) * br + (y - dy * r0)) * br + (x - dx * r0)];
                }
        brick[(z * br + y) * br + x] += acc;
    }
}

// 8^3 bricks with apron 1: the 296 interior voxels on the brick's boundary
// layer (the only ones copied into a neighbour's apron), ascending.
struct AdjTab10 {
    uint16_t i[296];
};
constexpr AdjTab10 make_adj10() {
    AdjTab10 t{};
    int n = 0;
    for (int i = 0; i < 1000; i++) {
        const int x = i % 10, y = (i / 10) % 10, z = i / 100;
        const bool in = x >= 1 && x <= 8 && y >= 1 && y <= 8 && z >= 1 && z <= 8;
        if (in && (x == 1 || x == 8 || y == 1 || y == 8 || z == 1 || z == 8)) t.i[n++] = (uint16_t)i;
    }
    return t;
}
__constant__ AdjTab10 ADJ10 = make_adj10();

// One wave per brick, XCD-contiguous runs of bricks; each lane's (up to 5)
// voxels issue all their neighbour reads before summing.  Same terms in the
// same order (dz, dy, dx ascending) as k_apron_adjoint.
__global__ __launch_bounds__(256) void k_apron_adjoint10(ApronParams ap, float* grad, int64_t nblk) {
    const int lane = threadIdx.x & 63;
    const int64_t blk = gvdbx_xcd_item(blockIdx.x, nblk);
    if (blk < 0) return;
    const int64_t k = blk * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (k >= ap.nleaf) return;
    const GNode nd = ap.leaves[k];
    int32_t leaf = -1;
    if (lane < 27) {
        const int cx = nd.mPos[0] >> 3, cy = nd.mPos[1] >> 3, cz = nd.mPos[2] >> 3;
        const int nx = cx + lane % 3 - 1, ny = cy + (lane / 3) % 3 - 1, nz = cz + lane / 9 - 1;
        const int64_t ncell = (nx >= 0 && ny >= 0 && nz >= 0) ? ap.g0.find(nx, ny, nz) : -1;
        if (ncell >= 0) leaf = ap.idx0[ncell];
        if (leaf >= ap.nleaf) leaf = -1;
    }
    int nbl[27];
#pragma unroll
    for (int d = 0; d < 27; d++) nbl[d] = __builtin_amdgcn_readlane(leaf, d);
    float* brick = grad + k * 1000;
#pragma unroll
    for (int u = 0; u < 5; u++) {
        const int j = lane + 64 * u;
        if (j >= 296) break;
        const int i = ADJ10.i[j];
        const int x = i % 10, y = (i / 10) % 10, z = i / 100;
        const int lx = x == 1 ? -1 : 0, hx = x == 8 ? 1 : 0;
        const int ly = y == 1 ? -1 : 0, hy = y == 8 ? 1 : 0;
        const int lz = z == 1 ? -1 : 0, hz = z == 8 ? 1 : 0;
        float v[27];
        bool on[27];
#pragma unroll
        for (int d = 0; d < 27; d++) {
            const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
            on[d] = d != 13 && dx >= lx && dx <= hx && dy >= ly && dy <= hy && dz >= lz && dz <= hz && nbl[d] >= 0;
            v[d] = on[d] ? grad[(int64_t)nbl[d] * 1000 + i - 8 * (dz * 100 + dy * 10 + dx)] : 0.0f;
        }
        float acc = 0.0f;
#pragma unroll
        for (int d = 0; d < 27; d++)
            if (on[d]) acc += v[d];
        brick[i] += acc;
    }
}

// Adjoint of the cubic splat, one thread per bin-sorted point:
//   dL/dw = sum_v g(v) t^3,  dL/dp = sum_v g(v) w 6 t^2 (v - p) / r^2,
//   t = 1 - |v-p|^2/r^2, over integer voxels v of active leaves, |v-p| < r.
__global__ __launch_bounds__(256) void k_splat_backward(SplatBackParams p) {
    const int64_t total = *p.total;
    const int br = p.brick_res;
    const int64_t nv = (int64_t)br * br * br;
    const float r2 = p.radius * p.radius;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total && j < p.nmax;
         j += (int64_t)gridDim.x * blockDim.x) {
        const float4 q = p.sorted[j];
        const int i = __float_as_int(q.w);
        int sh[3] = {0, 0, 0};   // cloud shift in level-0 cells
        if (p.cm.ncloud > 0) {
            const int cl = p.cm.cloud_of_point(i);
#pragma unroll
            for (int a = 0; a < 3; a++) sh[a] = p.cm.cell_off[3 * cl + a];
        }
        const float wt = p.weights ? *(const float*)(p.weights + (int64_t)i * p.wstride) : 1.0f;
        const int x0 = max((int)floorf(q.x - p.radius), 0), x1 = (int)ceilf(q.x + p.radius);
        const int y0 = max((int)floorf(q.y - p.radius), 0), y1 = (int)ceilf(q.y + p.radius);
        const int z0 = max((int)floorf(q.z - p.radius), 0), z1 = (int)ceilf(q.z + p.radius);
        float gw = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
        const float r2m = (p.radius + 0.01f) * (p.radius + 0.01f);   // row pruning (d2 < r2 decides)
        for (int z = z0; z <= z1; z++) {
            const float dz = (float)z - q.z;
            const float dz2 = dz * dz;
            if (dz2 > r2m) continue;
            for (int y = y0; y <= y1; y++) {
                const float dy = (float)y - q.y;
                const float row = dz2 + dy * dy;
                if (row > r2m) continue;
                const float half = sqrtf(r2m - row);
                const int xa = max(x0, (int)ceilf(q.x - half)), xb = min(x1, (int)floorf(q.x + half));
                for (int x = xa; x <= xb; x++) {
                    const float dx = (float)x - q.x;
                    const float d2 = gvdbx_d2(dx, dy, dz);
                    if (!(d2 < r2)) continue;
                    const int cx = x >> p.logr0, cy = y >> p.logr0, cz = z >> p.logr0;
                    const int64_t lcell = p.g0.find(cx + sh[0], cy + sh[1], cz + sh[2]);
                    if (lcell < 0) continue;
                    const int32_t leaf = p.idx0[lcell];
                    if (leaf < 0 || leaf >= p.nleaf) continue;
                    const int lx = x - (cx << p.logr0) + 1, ly = y - (cy << p.logr0) + 1, lz = z - (cz << p.logr0) + 1;
                    const float g = p.grad[(int64_t)leaf * nv + (lz * br + ly) * br + lx];
                    const float t = 1.0f - d2 / r2;
                    gw += g * (t * t * t);
                    const float f = g * wt * 6.0f * (t * t) / r2;
                    gx += f * dx;
                    gy += f * dy;
                    gz += f * dz;
                }
            }
        }
        if (p.dw) p.dw[i] = gw;
        p.dpos[3 * (int64_t)i] = gx;
        p.dpos[3 * (int64_t)i + 1] = gy;
        p.dpos[3 * (int64_t)i + 2] = gz;
    }
}

inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace

void launch_render(int shade, RenderTree tv, RenderCam cam, RenderScene sc, float4* rgba, uchar4* rgba8,
                   hipStream_t s) {
    dim3 grid((unsigned)((cam.width + 15) / 16), (unsigned)((cam.height + 15) / 16));
    const bool b10 = tv.brick_res == 10;
#define GVDBX_RENDER(K)                                                                          \
    do {                                                                                         \
        if (b10) hipLaunchKernelGGL(K<10>, grid, dim3(256), 0, s, tv, cam, sc, rgba, rgba8);    \
        else hipLaunchKernelGGL(K<0>, grid, dim3(256), 0, s, tv, cam, sc, rgba, rgba8);          \
    } while (0)
#define GVDBX_RENDER_SURF(M)                                                                                    \
    do {                                                                                                        \
        if (b10) hipLaunchKernelGGL((k_render_surface<M, 10>), grid, dim3(256), 0, s, tv, cam, sc, rgba, rgba8); \
        else hipLaunchKernelGGL((k_render_surface<M, 0>), grid, dim3(256), 0, s, tv, cam, sc, rgba, rgba8);      \
    } while (0)
    switch (shade) {
        case GVDBX_SHADE_VOLUME: GVDBX_RENDER(k_render_volume); break;
        case GVDBX_SHADE_TRILINEAR: GVDBX_RENDER_SURF(MODE_SURF_TRI); break;
        case GVDBX_SHADE_LEVELSET: GVDBX_RENDER_SURF(MODE_SURF_LS); break;
        case GVDBX_SHADE_VOXEL: GVDBX_RENDER_SURF(MODE_SURF_VOXEL); break;
        case GVDBX_SHADE_TRICUBIC: GVDBX_RENDER_SURF(MODE_SURF_CUBIC); break;
        case GVDBX_SHADE_EMPTYSKIP: GVDBX_RENDER(k_render_empty); break;
        case GVDBX_SHADE_SECTION2D: GVDBX_RENDER(k_render_section2d); break;
        case GVDBX_SHADE_SECTION3D: GVDBX_RENDER(k_render_section3d); break;
        default: {
            const int64_t n = (int64_t)cam.width * cam.height;
            hipLaunchKernelGGL(k_render_off, dim3(grid_for(n, 256)), dim3(256), 0, s, n, rgba, rgba8);
        }
    }
#undef GVDBX_RENDER
#undef GVDBX_RENDER_SURF
}

void launch_raytrace(RenderTree tv, RenderScene sc, GScnRay* rays, int64_t n, float bias, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (tv.brick_res == 10) hipLaunchKernelGGL(k_raytrace<10>, dim3(g), dim3(256), 0, s, tv, sc, rays, n, bias);
    else hipLaunchKernelGGL(k_raytrace<0>, dim3(g), dim3(256), 0, s, tv, sc, rays, n, bias);
}

void launch_raycast_ea(RenderTree tv, RenderCam cam, const RenderCam* cams, int ncam, RenderScene sc, float* I,
                       float* T, int record, VisitRec vr, hipStream_t s) {
    dim3 grid((unsigned)((cam.width + 15) / 16), (unsigned)((cam.height + 15) / 16), (unsigned)(cams ? ncam : 1));
    const bool b10 = tv.brick_res == 10;
    if (record) {
        if (b10) hipLaunchKernelGGL((k_raycast_ea<10, true>), grid, dim3(256), 0, s, tv, cam, cams, sc, I, T, vr);
        else hipLaunchKernelGGL((k_raycast_ea<0, true>), grid, dim3(256), 0, s, tv, cam, cams, sc, I, T, vr);
    } else {
        if (b10) hipLaunchKernelGGL((k_raycast_ea<10, false>), grid, dim3(256), 0, s, tv, cam, cams, sc, I, T, vr);
        else hipLaunchKernelGGL((k_raycast_ea<0, false>), grid, dim3(256), 0, s, tv, cam, cams, sc, I, T, vr);
    }
}

void launch_segments(const uint32_t* key, int64_t n, int32_t* seg, int64_t nleaf, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_segments, dim3(grid_for(n, 256)), dim3(256), 0, s, key, n, seg, nleaf);
}

void launch_chunk_order(const uint64_t* tag, int64_t nchunk, const uint32_t* chunk_base, const uint32_t* chunk_size,
                        const int32_t* wave_off, const uint32_t* key, uint32_t* key_out, uint32_t* val_out,
                        hipStream_t s) {
    if (nchunk > 0)
        hipLaunchKernelGGL(k_chunk_order, dim3((unsigned)nchunk), dim3(64), 0, s, tag, nchunk, chunk_base, chunk_size,
                           wave_off, key, key_out, val_out);
}

void launch_ea_backward(EABackParams p, hipStream_t s) {
    // the padded brick is staged in LDS; only the 10^3 brick (8^3 + apron 1)
    // has a kernel, checked by the caller
    if (p.tv.ncnt[0] > 0)
        hipLaunchKernelGGL(k_ea_backward<10>, dim3((unsigned)p.tv.ncnt[0]), dim3(256), 0, s, p);
}

void launch_apron_adjoint(ApronParams ap, float* grad, hipStream_t s) {
    if (ap.nleaf <= 0) return;
    if (ap.brick_res == 10 && ap.range0 == 8)
        hipLaunchKernelGGL(k_apron_adjoint10, dim3((unsigned)gvdbx_xcd_blocks((ap.nleaf + 3) / 4)), dim3(256), 0, s, ap,
                           grad, (ap.nleaf + 3) / 4);
    else if (ap.brick_res == 10)
        hipLaunchKernelGGL(k_apron_adjoint<10>, dim3((unsigned)ap.nleaf), dim3(256), 0, s, ap, grad);
    else
        hipLaunchKernelGGL(k_apron_adjoint<0>, dim3((unsigned)ap.nleaf), dim3(256), 0, s, ap, grad);
}

void launch_splat_backward(SplatBackParams p, hipStream_t s) {
    if (p.nmax > 0) hipLaunchKernelGGL(k_splat_backward, dim3(grid_for(p.nmax, 256)), dim3(256), 0, s, p);
}
