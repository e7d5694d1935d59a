// VolumeGVDB.h -- header-only C++ facade over the gvdbx C-ABI.
//
// Keeps the method names and argument meaning of nvdb::VolumeGVDB
// (source/gvdb_library/src/gvdb_volume_gvdb.h:319-901) for the points path,
// so the reference samples' call sequences (e.g.
// source/gFluidSurface/main_fluid_surface.cpp:320-356) port line for line:
//
//   gvdbx::VolumeGVDB gvdb;                    // SetCudaDevice + Initialize
//   gvdb.Configure(3, 3, 3, 3, 5);
//   gvdb.SetChannelDefault(16, 16, 1);
//   gvdb.AddChannel(0, T_FLOAT, 1);
//   gvdb.SetDataGPU(pntpos, n, dev_ptr, 0, sizeof(float) * 3);
//   gvdb.SetPoints(pntpos, pntvel, pntclr);
//   gvdb.RebuildTopology(n, radius * 2.0f, origin);
//   gvdb.FinishTopology(false, true);
//   gvdb.UpdateAtlas();
//   int scPntLen = 0;
//   gvdb.InsertPointsSubcell(4, n, radius * 2.0f, origin, scPntLen);
//   gvdb.GatherLevelSet(4, n, radius, origin, scPntLen, 0, 1);
//   gvdb.UpdateApron(0, 3.0f);
//
// Differences from the reference: errors throw gvdbx::Error (the reference
// prints and exits); the DataPtr handles are non-owning device views.
#pragma once
#include <climits>
#include <stdexcept>
#include <string>
#include <vector>

#include "gvdbx/gvdbx_c.h"

namespace gvdbx {

struct Vector3DF {
    float x = 0, y = 0, z = 0;
    Vector3DF() = default;
    Vector3DF(float a, float b, float c) : x(a), y(b), z(c) {}
};

// Non-owning device buffer view (reference DataPtr, gvdb_allocator.h:35-55).
struct DataPtr {
    const void* gpu = nullptr;
    int lastEle = 0;
    int offset = 0;
    int stride = 0;
};

enum { T_UCHAR4 = GVDBX_UCHAR4, T_FLOAT = GVDBX_FLOAT };
enum { F_LINEAR = 1, F_CLAMP = 1 };

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int code, const char* fn) {
    if (code != GVDBX_OK) throw Error(code, std::string(fn) + ": " + gvdbx_last_error());
}

class VolumeGVDB {
public:
    explicit VolumeGVDB(int device = 0) { check(gvdbx_create(&v_, device), "gvdbx_create"); }
    ~VolumeGVDB() {
        if (v_) gvdbx_destroy(v_);
    }
    VolumeGVDB(const VolumeGVDB&) = delete;
    VolumeGVDB& operator=(const VolumeGVDB&) = delete;

    gvdbx_volume* handle() const { return v_; }
    void SetStream(void* hip_stream) { check(gvdbx_set_stream(v_, hip_stream), "SetStream"); }

    // Configure(q4,q3,q2,q1,q0): level 0 (the brick) takes q0 (VG.h:387)
    void Configure(int q4, int q3, int q2, int q1, int q0) {
        int r[5] = {q0, q1, q2, q3, q4};
        check(gvdbx_configure(v_, 5, r), "Configure");
    }
    // Configure(levs, r, numcnt) (VG.h:388); numcnt is accepted for signature
    // compatibility: pools grow on demand
    void Configure(int levs, int* r, int* numcnt, bool use_masks = false) {
        (void)numcnt;
        (void)use_masks;
        check(gvdbx_configure(v_, levs, r), "Configure");
    }
    void SetChannelDefault(int cx, int cy, int cz) { check(gvdbx_set_channel_default(v_, cx, cy, cz), "SetChannelDefault"); }
    void AddChannel(unsigned char chan, int dt, int apron, int filter = F_LINEAR, int border = F_CLAMP) {
        check(gvdbx_add_channel(v_, chan, dt, apron, filter, border), "AddChannel");
    }
    void FillChannel(unsigned char chan, float value) { check(gvdbx_fill_channel(v_, chan, value), "FillChannel"); }

    // SetDataGPU (VG.cpp:4872-4880): wrap caller device memory, no copy
    void SetDataGPU(DataPtr& p, int cnt, const void* gpubuf, int offs, int stride) {
        p.gpu = gpubuf;
        p.lastEle = cnt;
        p.offset = offs;
        p.stride = stride;
    }
    // SetPoints(pos, vel, clr) (VG.h:541)
    void SetPoints(DataPtr& pnts, DataPtr& vel, DataPtr& clr) {
        check(gvdbx_set_points(v_, pnts.gpu, pnts.offset, pnts.stride ? pnts.stride : 12, vel.gpu, vel.offset,
                               vel.stride, clr.gpu, clr.offset, clr.stride),
              "SetPoints");
    }
    void SetPointWeights(const float* w_dev, int stride = 4) {
        check(gvdbx_set_point_weights(v_, w_dev, stride), "SetPointWeights");
    }

    void RebuildTopology(int num_pnts, float radius, Vector3DF orig) {
        check(gvdbx_rebuild_topology(v_, num_pnts, radius, orig.x, orig.y, orig.z), "RebuildTopology");
    }
    void AccumulateTopology(int num_pnts, float radius, Vector3DF orig, int iDepth = 1) {
        check(gvdbx_accumulate_topology(v_, num_pnts, radius, orig.x, orig.y, orig.z, iDepth), "AccumulateTopology");
    }
    void RequestFullRebuild(bool tf) { check(gvdbx_request_full_rebuild(v_, tf ? 1 : 0), "RequestFullRebuild"); }
    void SaveVBX(const std::string& fname) { check(gvdbx_save_vbx(v_, fname.c_str()), "SaveVBX"); }
    bool LoadVBX(const std::string& fname, int force_maj = 0, int force_min = 0) {
        check(gvdbx_load_vbx_version(v_, fname.c_str(), force_maj, force_min), "LoadVBX");
        return true;
    }
    void SetTransform(Vector3DF pretrans, Vector3DF scal, Vector3DF angs, Vector3DF trans) {
        const float p[3] = {pretrans.x, pretrans.y, pretrans.z}, s[3] = {scal.x, scal.y, scal.z};
        const float a[3] = {angs.x, angs.y, angs.z}, t[3] = {trans.x, trans.y, trans.z};
        check(gvdbx_set_transform(v_, p, s, a, t), "SetTransform");
    }
    void FinishTopology(bool commit = false, bool compute_bounds = true) {
        check(gvdbx_finish_topology(v_, commit, compute_bounds), "FinishTopology");
    }
    void UpdateAtlas() { check(gvdbx_update_atlas(v_), "UpdateAtlas"); }

    // the reference's out-parameter is an int: a list longer than INT_MAX
    // throws GVDBX_ERR_RANGE here (the int64_t overload returns any length)
    void InsertPointsSubcell(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength) {
        int64_t len = 0;
        InsertPointsSubcell(subcell_size, num_pnts, radius, trans, len);
        if (len > INT_MAX)
            throw Error(GVDBX_ERR_RANGE, "InsertPointsSubcell: list length " + std::to_string(len) +
                                             " does not fit the int out-parameter (use the int64_t overload)");
        pSCPntsLength = (int)len;
    }
    void InsertPointsSubcell(int subcell_size, int num_pnts, float radius, Vector3DF trans, int64_t& pSCPntsLength) {
        check(gvdbx_insert_points_subcell(v_, subcell_size, num_pnts, radius, trans.x, trans.y, trans.z,
                                          &pSCPntsLength),
              "InsertPointsSubcell");
    }
    void InsertPointsSubcell_FP16(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength) {
        int64_t len = 0;
        check(gvdbx_insert_points_subcell_fp16(v_, subcell_size, num_pnts, radius, trans.x, trans.y, trans.z, &len),
              "InsertPointsSubcell_FP16");
        if (len > INT_MAX) throw Error(GVDBX_ERR_RANGE, "InsertPointsSubcell_FP16: list length beyond int");
        pSCPntsLength = (int)len;
    }
    void GatherLevelSet_FP16(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                             int chanLevelset, int chanClr) {
        check(gvdbx_gather_level_set_fp16(v_, subcell_size, num_pnts, radius, trans.x, trans.y, trans.z,
                                          pSCPntsLength, chanLevelset, chanClr),
              "GatherLevelSet_FP16");
    }
    void GatherDensity(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                       int chanDensity, int chanClr, bool bAccum = false) {
        check(gvdbx_gather_density(v_, subcell_size, num_pnts, radius, trans.x, trans.y, trans.z, pSCPntsLength,
                                   chanDensity, chanClr, bAccum),
              "GatherDensity");
    }
    void GatherLevelSet(int subcell_size, int num_pnts, float radius, Vector3DF trans, int& pSCPntsLength,
                        int chanLevelset, int chanClr, bool bAccum = false) {
        check(gvdbx_gather_level_set(v_, subcell_size, num_pnts, radius, trans.x, trans.y, trans.z, pSCPntsLength,
                                     chanLevelset, chanClr, bAccum),
              "GatherLevelSet");
    }
    void InsertPoints(int num_pnts, Vector3DF trans, bool bPrefix = false) {
        check(gvdbx_insert_points(v_, num_pnts, trans.x, trans.y, trans.z, bPrefix), "InsertPoints");
    }
    void ScatterDensity(int num_pnts, float radius, float amp, Vector3DF trans, bool expand = true,
                        bool avgColor = false) {
        check(gvdbx_scatter_density(v_, num_pnts, radius, amp, trans.x, trans.y, trans.z, expand, avgColor),
              "ScatterDensity");
    }
    void SplatCubic(int num_pnts, float radius, Vector3DF trans, int chan = 0, bool accum = false) {
        check(gvdbx_splat_cubic(v_, num_pnts, radius, trans.x, trans.y, trans.z, chan, accum), "SplatCubic");
    }
    void UpdateApron(unsigned char chan, float boundval = 0.0f) {
        check(gvdbx_update_apron(v_, chan, boundval), "UpdateApron");
    }

    // ---- rendering (Render VG.cpp / rayCast RC.cuh with an explicit camera) ----
    void Render(int shade, unsigned char chan, const gvdbx_camera& cam, const gvdbx_render_params& rp,
                float* rgba_dev, uint8_t* rgba8_dev = nullptr) {
        check(gvdbx_render(v_, shade, chan, &cam, &rp, rgba_dev, rgba8_dev), "Render");
    }
    // SetColorChannel(uchar) VG.cpp:686-690: 255 (CHAN_UNDEF) for none
    void SetColorChannel(unsigned char chan) { check(gvdbx_set_color_channel(v_, chan), "SetColorChannel"); }
    void Raytrace(gvdbx_scn_ray* rays_dev, int64_t num_rays, unsigned char chan, const gvdbx_render_params& rp,
                  float bias) {
        check(gvdbx_raytrace(v_, chan, &rp, rays_dev, num_rays, bias), "Raytrace");
    }

    // ---- differentiable forms (no reference counterpart) ----
    void RaycastEA(unsigned char chan, const gvdbx_camera& cam, const gvdbx_render_params& rp, float* I_dev,
                   float* T_dev, bool record = false) {
        check(gvdbx_raycast_ea(v_, chan, &cam, &rp, I_dev, T_dev, record ? 1 : 0), "RaycastEA");
    }
    void RaycastEABackward(unsigned char chan, const gvdbx_camera& cam, const gvdbx_render_params& rp,
                           const float* I_dev, const float* T_dev, const float* gI_dev, const float* gT_dev,
                           float* grad_dev) {
        check(gvdbx_raycast_ea_backward(v_, chan, &cam, &rp, I_dev, T_dev, gI_dev, gT_dev, grad_dev),
              "RaycastEABackward");
    }
    void ApronAdjoint(float* grad_dev) { check(gvdbx_apron_adjoint(v_, grad_dev), "ApronAdjoint"); }
    void SplatCubicBackward(int num_pnts, float radius, Vector3DF trans, const float* grad_dev, float* dpos_dev,
                            float* dw_dev = nullptr) {
        check(gvdbx_splat_cubic_backward(v_, num_pnts, radius, trans.x, trans.y, trans.z, grad_dev, dpos_dev, dw_dev),
              "SplatCubicBackward");
    }

    // ---- batched clouds (no reference counterpart) ----
    void SetPointsBatched(const void* pos_dev, int stride, const std::vector<int64_t>& cloud_off) {
        check(gvdbx_set_points_batched(v_, pos_dev, stride, cloud_off.data(), (int)cloud_off.size() - 1),
              "SetPointsBatched");
    }
    void RaycastEABatched(unsigned char chan, const std::vector<gvdbx_camera>& cams, const gvdbx_render_params& rp,
                          float* I_dev, float* T_dev, bool record = false) {
        check(gvdbx_raycast_ea_batched(v_, chan, cams.data(), (int)cams.size(), &rp, I_dev, T_dev, record ? 1 : 0),
              "RaycastEABatched");
    }
    void RaycastEABackwardBatched(unsigned char chan, const std::vector<gvdbx_camera>& cams,
                                  const gvdbx_render_params& rp, const float* I_dev, const float* T_dev,
                                  const float* gI_dev, const float* gT_dev, float* grad_dev) {
        check(gvdbx_raycast_ea_backward_batched(v_, chan, cams.data(), (int)cams.size(), &rp, I_dev, T_dev, gI_dev,
                                                gT_dev, grad_dev),
              "RaycastEABackwardBatched");
    }

    // ---- kernel tuning (no reference counterpart) ----
    void SetTuning(const char* key, int value) { check(gvdbx_set_tuning(v_, key, value), "SetTuning"); }

    gvdbx_info Info() const {
        gvdbx_info inf;
        check(gvdbx_get_info(v_, &inf), "GetInfo");
        return inf;
    }
    // host copy of one channel: leaf-major bricks incl. apron
    std::vector<float> ReadBricks(int chan) const {
        gvdbx_info inf = Info();
        int64_t br = inf.brick_res;
        std::vector<float> out((size_t)(inf.node_cnt[0] * br * br * br));
        check(gvdbx_read_bricks(v_, chan, out.data(), (int64_t)out.size()), "ReadBricks");
        return out;
    }

private:
    gvdbx_volume* v_ = nullptr;
};

}  // namespace gvdbx
