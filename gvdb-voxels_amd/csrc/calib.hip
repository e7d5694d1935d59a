// calib.hip -- vector-issue calibration kernels for the benchmark (not part of
// the gvdbx C-ABI; built as libgvdbx_calib.so).
//
// Each kernel runs a fixed number of loop iterations per wave whose bodies
// issue a known number of wave64 VALU instructions with no memory traffic,
// so VALU instructions / elapsed time is the chip's vector issue rate for
// that instruction mix:
//   kind 0: eight independent FP32 FMA chains (8 VALU per iteration);
//   kind 1: the level-set walk's mix per candidate: three v_readlane of a
//           candidate's coordinates, two subtractions, a multiply and an
//           FMA for the column distance, then eight voxels of (subtract,
//           FMA, unsigned min) (3 + 4 + 24 = 31 VALU per iteration).
// The loop's own bookkeeping is scalar.  The result is written only when it
// equals a key passed at run time that it cannot equal, so the compiler
// keeps the work and nothing is stored.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ __launch_bounds__(256) void k_fma_mix(int iters, float seed, float key, float* sink) {
    float a0 = seed + threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3;
    float a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
    const float m = 0.999999f, c = 1e-7f;
    for (int i = 0; i < iters; i++) {
        a0 = __builtin_fmaf(a0, m, c);
        a1 = __builtin_fmaf(a1, m, c);
        a2 = __builtin_fmaf(a2, m, c);
        a3 = __builtin_fmaf(a3, m, c);
        a4 = __builtin_fmaf(a4, m, c);
        a5 = __builtin_fmaf(a5, m, c);
        a6 = __builtin_fmaf(a6, m, c);
        a7 = __builtin_fmaf(a7, m, c);
    }
    const float s = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
    if (s == key) sink[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_walk_mix(int iters, float seed, uint32_t key, uint32_t* sink) {
    const int lane = threadIdx.x & 63;
    const float vx = seed + lane, vy = seed - lane;
    float px = seed * 0.5f + lane, py = seed * 0.25f - lane, pz = seed + 0.125f * lane;
    uint32_t m0 = 0x7F800000u, m1 = m0, m2 = m0, m3 = m0, m4 = m0, m5 = m0, m6 = m0, m7 = m0;
    for (int i = 0; i < iters; i++) {
        const int b = i & 63;
        const float qx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(px), b));
        const float qy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(py), b));
        const float qz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pz), b));
        const float dx = vx - qx, dy = vy - qy;
        const float dxy = __builtin_fmaf(dy, dy, dx * dx);
        float dz;
        dz = vx - qz; m0 = min(m0, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = vy - qz; m1 = min(m1, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = px - qz; m2 = min(m2, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = py - qz; m3 = min(m3, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = qx - pz; m4 = min(m4, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = qy - pz; m5 = min(m5, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = qx - vy; m6 = min(m6, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
        dz = qy - vx; m7 = min(m7, __float_as_uint(__builtin_fmaf(dz, dz, dxy)));
    }
    const uint32_t s = min(min(min(m0, m1), min(m2, m3)), min(min(m4, m5), min(m6, m7)));
    if (s == key) sink[blockIdx.x] = s;
}

}  // namespace

extern "C" {

// Runs calibration kernel `kind` (0: FMA chains, 1: walk mix) on `blocks`
// workgroups of 256 threads, `iters` iterations per wave, on `stream`, and
// returns its duration (HIP events on that stream) and the wave64 VALU
// instructions its loop bodies issue.  Status 0 on success.
int gvdbx_calib_valu(int kind, int blocks, int iters, void* stream, double* ms_out, double* valu_out) {
    hipStream_t s = (hipStream_t)stream;
    if (blocks <= 0 || iters <= 0 || !ms_out || !valu_out || (kind != 0 && kind != 1)) return 1;
    void* sink = nullptr;
    if (hipMalloc(&sink, (size_t)blocks * 4) != hipSuccess) return 7;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto launch = [&]() {
        if (kind == 0)
            hipLaunchKernelGGL(k_fma_mix, dim3(blocks), dim3(256), 0, s, iters, 1.5f, -1.0f, (float*)sink);
        else
            hipLaunchKernelGGL(k_walk_mix, dim3(blocks), dim3(256), 0, s, iters, 1.5f, 0xFFFFFFFFu, (uint32_t*)sink);
    };
    launch();   // warm-up
    (void)hipEventRecord(e0, s);
    launch();
    (void)hipEventRecord(e1, s);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    if (err != hipSuccess) return 6;
    const double per_iter = kind == 0 ? 8.0 : 31.0;
    *ms_out = ms;
    *valu_out = (double)blocks * 4.0 * (double)iters * per_iter;   // 4 waves per workgroup
    return 0;
}

}  // extern "C"
