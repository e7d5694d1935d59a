// sort.hip -- stable LSD radix sort of (uint32 or uint64 key, uint32 value)
// pairs for gfx950.  Replaces the reference's Thrust sort (cuda_gvdb_copydata.cu:11-16,
// called from RadixSortByByte VG.cpp:904-928) wherever this library needs a
// sorted key list.
//
// One pass per 8-bit digit, three launches per pass:
//   1. k_rs_hist: per 4096-key tile, a 256-bin digit histogram in LDS,
//      written digit-major (hist[d * ntiles + tile]);
//   2. an exclusive scan over the digit-major histogram gives every
//      (digit, tile) its output base (the count scan of the topology code);
//   3. k_rs_scatter: each tile re-reads its keys in 16 rounds of 256 (index
//      order), ranks equal digits inside each wave with eight ballots, adds
//      the counts of lower waves and earlier rounds from LDS, reorders the
//      tile by digit in LDS, and writes every digit's run to its base.
//      Equal digits keep their input order: stable.
// Reads and writes 8 B (32-bit keys) or 12 B (64-bit keys) per pair per pass
// plus the histogram traffic.
#include "gvdbx_kernels.h"

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;

inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K* keys, int64_t n, int shift, int64_t ntiles,
                                                         int32_t* hist) {
    __shared__ int32_t cnt[RS_BINS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t i = t0 + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFF], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// The tile is first reordered by digit in LDS (stable), then written out in
// that order: consecutive threads write consecutive addresses of one digit's
// run, so the global stores coalesce.
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K* keys, const uint32_t* vals, int64_t n,
                                                            int shift, int64_t ntiles, const int32_t* hist,
                                                            const int32_t* off, K* okeys, uint32_t* ovals) {
    __shared__ int32_t gbase[RS_BINS];    // global output start of each digit for this tile
    __shared__ int32_t tstart[RS_BINS];   // start of each digit inside the tile
    __shared__ int32_t run[RS_BINS];
    __shared__ int32_t wcnt[RS_WAVES][RS_BINS];
    __shared__ int32_t wtot[RS_WAVES];
    __shared__ K lkey[RS_TILE];
    __shared__ uint32_t lval[RS_TILE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * RS_TILE;
    const int tn = (int)min((int64_t)RS_TILE, n - t0);
    gbase[tid] = off[(int64_t)tid * ntiles + blockIdx.x];
    run[tid] = 0;
    for (int w = 0; w < RS_WAVES; w++) wcnt[w][tid] = 0;
    // exclusive scan of this tile's digit counts (from the histogram pass)
    {
        const int c = hist[(int64_t)tid * ntiles + blockIdx.x];
        int x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wtot[wave] = x;
        __syncthreads();
        int pre = 0;
        for (int w = 0; w < wave; w++) pre += wtot[w];
        tstart[tid] = pre + x - c;
        __syncthreads();
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < RS_ITEMS; r++) {
        const int li = r * RS_THREADS + tid;
        const bool ok = li < tn;
        K k = 0;
        uint32_t val = 0;
        int d = 0;
        if (ok) {
            k = keys[t0 + li];
            val = vals[t0 + li];
            d = (int)((k >> shift) & 0xFF);
        }
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot(ok && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        if (ok && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (ok) {
            int pos = tstart[d] + run[d] + rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][d];
            lkey[pos] = k;
            lval[pos] = val;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < RS_WAVES; w++) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
    for (int j = tid; j < tn; j += RS_THREADS) {
        const K k = lkey[j];
        const int d = (int)((k >> shift) & 0xFF);
        const int64_t pos = (int64_t)gbase[d] + (j - tstart[d]);
        okeys[pos] = k;
        ovals[pos] = lval[j];
    }
}

__global__ void k_iota_u32(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

}  // namespace

int64_t radix_hist_entries(int64_t n) {
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    if (ntiles < 1) ntiles = 1;
    return ntiles * RS_BINS;
}

void launch_iota_u32(uint32_t* p, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
}

template <typename K>
int radix_sort_impl(K* keys, uint32_t* vals, K* tkeys, uint32_t* tvals, int64_t n, int begin_bit, int end_bit,
                    int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s) {
    if (n <= 1 || end_bit <= begin_bit) return 0;
    const int64_t ntiles = radix_hist_entries(n) / RS_BINS;
    const int64_t nh = ntiles * RS_BINS;
    const int64_t stiles = scan_tiles(nh + 1);
    K* kin = keys;
    uint32_t* vin = vals;
    K* kout = tkeys;
    uint32_t* vout = tvals;
    int flips = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, kin, n, shift, ntiles, hist);
        launch_count_scan_reduce(hist, nh, partials, partials + stiles, s);
        launch_count_scan_apply(hist, nh, partials, off, s);
        hipLaunchKernelGGL(k_rs_scatter<K>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, kin, vin, n, shift,
                           ntiles, hist, off, kout, vout);
        K* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
        flips ^= 1;
    }
    return flips;
}

int radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* tkeys, uint32_t* tvals, int64_t n, int begin_bit,
                     int end_bit, int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s) {
    return radix_sort_impl<uint64_t>(keys, vals, tkeys, tvals, n, begin_bit, end_bit, hist, off, partials, s);
}

int radix_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t* tkeys, uint32_t* tvals, int64_t n, int begin_bit,
                       int end_bit, int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s) {
    return radix_sort_impl<uint32_t>(keys, vals, tkeys, tvals, n, begin_bit, end_bit, hist, off, partials, s);
}
