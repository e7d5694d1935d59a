// batch.hip -- batched inputs: many depth frames (or many point clouds) are
// turned into one point set so that every later stage runs once over all of
// them (one launch per stage instead of one per frame or cloud).
#include "gvdbx/gvdbx_c.h"
#include "gvdbx_kernels.h"

namespace {

inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// One thread per pixel of every frame; the frame's pose is read once per
// thread (12 floats, shared by the frame's pixels through the cache).
__global__ __launch_bounds__(256) void k_backproject(const float* depth, int64_t total, int width, int height,
                                                     float fx, float fy, float cx, float cy, const float* poses,
                                                     float voxel, float ox, float oy, float oz, float dmin,
                                                     float dmax, float* out) {
    const int64_t per = (int64_t)width * height;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = i / per;
        const int64_t pix = i - f * per;
        const int v = (int)(pix / width), u = (int)(pix - (int64_t)v * width);
        const float d = depth[i];
        float px = 0.0f, py = 0.0f, pz = GVDBX_NOHIT;
        if (d >= dmin && d <= dmax) {   // false for NaN
            const float* P = poses + 12 * f;
            const float c0 = d * (((float)u - cx) / fx);
            const float c1 = d * (((float)v - cy) / fy);
            const float c2 = d;
            const float w0 = ((P[0] * c0 + P[1] * c1) + P[2] * c2) + P[9];
            const float w1 = ((P[3] * c0 + P[4] * c1) + P[5] * c2) + P[10];
            const float w2 = ((P[6] * c0 + P[7] * c1) + P[8] * c2) + P[11];
            px = (w0 - ox) / voxel;
            py = (w1 - oy) / voxel;
            pz = (w2 - oz) / voxel;
        }
        out[3 * i] = px;
        out[3 * i + 1] = py;
        out[3 * i + 2] = pz;
    }
}

}  // namespace

void launch_backproject(const float* depth, int nframes, int width, int height, const float* intr4,
                        const float* poses, float voxel, const float* origin3, float dmin, float dmax, float* out,
                        hipStream_t s) {
    const int64_t total = (int64_t)nframes * width * height;
    if (total > 0)
        hipLaunchKernelGGL(k_backproject, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0, s, depth, total, width,
                           height, intr4[0], intr4[1], intr4[2], intr4[3], poses, voxel, origin3[0], origin3[1],
                           origin3[2], dmin, dmax, out);
}
