// sort.hip -- stable LSD radix sort of (uint32 or uint64 key, uint32 value)
// pairs for gfx950.  Replaces the reference's Thrust sort (cuda_gvdb_copydata.cu:11-16,
// called from RadixSortByByte VG.cpp:904-928) wherever this library needs a
// sorted key list.
//
// The key bits [begin, end) are split into as few passes as digits of at
// most 10 bits (9 for 64-bit keys, whose tile takes more LDS) allow, of
// near-equal width (18 bits: two 9-bit passes).  Three launches per pass:
//   1. k_rs_hist: per 4096-key tile, a 2^DB-bin digit histogram in LDS,
//      written digit-major (hist[d * ntiles + tile]);
//   2. an exclusive scan over the digit-major histogram gives every
//      (digit, tile) its output base (the count scan of the topology code);
//   3. k_rs_scatter: each wave ranks a quarter of the tile (contiguous keys,
//      64 per round, DB ballots per round, its own running digit counts in
//      LDS: no block barrier between rounds), the per-wave counts become
//      offsets, the tile is reordered by digit in LDS and every digit's run
//      is written to its base.  Equal digits keep their input order: stable.
// Reads and writes 8 B (32-bit keys) or 12 B (64-bit keys) per pair per pass
// plus the histogram traffic.
#include "gvdbx_kernels.h"

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_MAX_BITS = 10;
constexpr int RS_MAX_BINS = 1 << RS_MAX_BITS;
constexpr int RS_WAVES = RS_THREADS / 64;

inline unsigned grid_for(int64_t n, int threads, int64_t cap = 65536) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// DB: digit bits of the kernel (2^DB bins); the pass's digit is
// (key >> shift) & mask with mask < 2^DB.
template <typename K, int DB, int ITEMS = RS_ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K* keys, int64_t n, int shift, uint32_t mask,
                                                         int64_t ntiles, int32_t* hist) {
    constexpr int B = 1 << DB;
    __shared__ int32_t cnt[B];
    for (int b = threadIdx.x; b < B; b += RS_THREADS) cnt[b] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * RS_THREADS * ITEMS;
    K kr[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {   // all loads first, then the LDS counts
        const int64_t i = t0 + r * RS_THREADS + threadIdx.x;
        kr[r] = i < n ? keys[i] : (K)0;
    }
    // Keys often come in runs (segments of neighbouring rays, points near in
    // index): lanes holding equal digits would all hit one LDS word, and
    // such atomics serialize.  Each run of equal digits over consecutive
    // lanes adds its length once, from its first lane.
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const bool ok = t0 + r * RS_THREADS + threadIdx.x < n;
        const uint32_t d = (uint32_t)(kr[r] >> shift) & mask;
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1);
        const uint64_t okm = __ballot(ok);
        const bool head = ok && (lane == 0 || d != prev);
        const uint64_t heads = __ballot(head);
        if (head) {
            // the run ends at the next head or at the last valid lane (valid
            // lanes are a prefix of the wave)
            const uint64_t after = heads & ~((2ull << lane) - 1ull);
            const int end = after ? __ffsll((unsigned long long)after) - 1 : __popcll(okm);
            atomicAdd(&cnt[d], end - lane);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += RS_THREADS) hist[(int64_t)b * ntiles + blockIdx.x] = cnt[b];
}

// The tile is first reordered by digit in LDS (stable), then written out in
// that order: consecutive threads write consecutive addresses of one digit's
// run, so the global stores coalesce.  Thread t owns bins [t R, t R + R).
template <typename K, int DB, int ITEMS = RS_ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K* keys, const uint32_t* vals, int64_t n,
                                                            int shift, uint32_t mask, int64_t ntiles,
                                                            const int32_t* hist, const int32_t* off, K* okeys,
                                                            uint32_t* ovals) {
    constexpr int B = 1 << DB;
    constexpr int R = B / RS_THREADS > 0 ? B / RS_THREADS : 1;
    static_assert(B >= RS_THREADS, "one bin or more per thread");
    __shared__ int32_t gbase[B];    // global output start of each digit for this tile
    __shared__ int32_t tstart[B];   // start of each digit inside the tile
    __shared__ int32_t wcnt[RS_WAVES][B];   // per wave: running digit counts, then offsets
    __shared__ int32_t wtot[RS_WAVES];
    __shared__ K lkey[(RS_THREADS * ITEMS)];
    __shared__ uint32_t lval[(RS_THREADS * ITEMS)];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * (RS_THREADS * ITEMS);
    const int tn = (int)min((int64_t)(RS_THREADS * ITEMS), n - t0);
    int c[R];
    int csum = 0;
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int b = tid * R + j;
        gbase[b] = off[(int64_t)b * ntiles + blockIdx.x];
        for (int w = 0; w < RS_WAVES; w++) wcnt[w][b] = 0;
        c[j] = hist[(int64_t)b * ntiles + blockIdx.x];
        csum += c[j];
    }
    // exclusive scan of this tile's digit counts (from the histogram pass)
    {
        int x = csum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wtot[wave] = x;
        __syncthreads();
        int pre = 0;
        for (int w = 0; w < wave; w++) pre += wtot[w];
        int st = pre + x - csum;
#pragma unroll
        for (int j = 0; j < R; j++) {
            tstart[tid * R + j] = st;
            st += c[j];
        }
        __syncthreads();
    }
    // Wave w ranks the tile's keys [w * 64 * ITEMS, (w + 1) * 64 * ITEMS)
    // on its own: 64 per round, lanes in order, against its running digit
    // counts wcnt[w] (only the wave touches them, and its LDS accesses stay
    // in program order), so the rounds need no block barrier.  Then the
    // counts become per-wave offsets (exclusive over waves) and every key
    // goes to tstart[d] + offset of its wave + its rank: equal digits keep
    // their tile order.
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int wbase = wave * 64 * ITEMS;
    K kr[ITEMS];
    uint32_t vr[ITEMS];
    int rk[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {   // all loads in flight at once
        const int li = wbase + r * 64 + lane;
        kr[r] = 0;
        vr[r] = 0;
        if (li < tn) {
            kr[r] = keys[t0 + li];
            vr[r] = vals ? vals[t0 + li] : (uint32_t)(t0 + li);   // no vals: each entry's index
        }
    }
    int32_t* wc = wcnt[wave];
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const bool ok = wbase + r * 64 + lane < tn;
        const int d = ok ? (int)((uint32_t)(kr[r] >> shift) & mask) : 0;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            uint64_t m = __ballot(ok && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        const int before = ok ? wc[d] : 0;
        rk[r] = before + rank;
        if (ok && rank == 0) wc[d] = before + __popcll(peers);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int b = tid * R + j;
        int acc = 0;
        for (int w = 0; w < RS_WAVES; w++) {
            const int t = wcnt[w][b];
            wcnt[w][b] = acc;
            acc += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        if (wbase + r * 64 + lane < tn) {
            const int d = (int)((uint32_t)(kr[r] >> shift) & mask);
            const int pos = tstart[d] + wc[d] + rk[r];
            lkey[pos] = kr[r];
            lval[pos] = vr[r];
        }
    }
    __syncthreads();
    for (int j = tid; j < tn; j += RS_THREADS) {
        const K k = lkey[j];
        const int d = (int)((uint32_t)(k >> shift) & mask);
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
    return ntiles * RS_MAX_BINS;
}

void launch_iota_u32(uint32_t* p, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
}

template <typename K, int DB, int ITEMS>
void radix_pass(const K* kin, const uint32_t* vin, K* kout, uint32_t* vout, int64_t n, int shift, int width,
                int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s) {
    const int64_t ntiles = std::max<int64_t>((n + RS_THREADS * ITEMS - 1) / (RS_THREADS * ITEMS), 1);
    const int64_t nh = ntiles << DB;
    const int64_t stiles = scan_tiles(nh + 1);
    const uint32_t mask = (1u << width) - 1u;
    hipLaunchKernelGGL((k_rs_hist<K, DB, ITEMS>), dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, kin, n, shift, mask,
                       ntiles, hist);
    launch_count_scan_reduce(hist, nh, partials, partials + stiles, s);
    launch_count_scan_apply(hist, nh, partials, off, s);
    hipLaunchKernelGGL((k_rs_scatter<K, DB, ITEMS>), dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, kin, vin, n, shift,
                       mask, ntiles, hist, off, kout, vout);
}

template <typename K>
int radix_sort_impl(K* keys, uint32_t* vals, K* tkeys, uint32_t* tvals, int64_t n, int begin_bit, int end_bit,
                    int32_t* hist, int32_t* off, uint32_t* partials, bool iota, bool wide, hipStream_t s) {
    if (n <= 1 || end_bit <= begin_bit) {
        if (iota) launch_iota_u32(vals, n, s);
        return 0;
    }
    const int maxdb = sizeof(K) == 8 ? 9 : RS_MAX_BITS;
    const int bits = end_bit - begin_bit;
    const int np = (bits + maxdb - 1) / maxdb;
    K* kin = keys;
    uint32_t* vin = vals;
    K* kout = tkeys;
    uint32_t* vout = tvals;
    int flips = 0;
    for (int p = 0, shift = begin_bit; p < np; p++) {
        const int width = bits / np + (p < bits % np ? 1 : 0);
        const uint32_t* vsrc = iota && p == 0 ? nullptr : vin;   // the first pass makes the indices itself
        if (wide) {
            if (width <= 8)
                radix_pass<K, 8, 2 * RS_ITEMS>(kin, vsrc, kout, vout, n, shift, width, hist, off, partials, s);
            else if (width == 9)
                radix_pass<K, 9, 2 * RS_ITEMS>(kin, vsrc, kout, vout, n, shift, width, hist, off, partials, s);
            else if constexpr (sizeof(K) == 4)
                radix_pass<K, 10, 2 * RS_ITEMS>(kin, vsrc, kout, vout, n, shift, width, hist, off, partials, s);
        } else {
            if (width <= 8)
                radix_pass<K, 8, RS_ITEMS>(kin, vsrc, kout, vout, n, shift, width, hist, off, partials, s);
            else if (width == 9)
                radix_pass<K, 9, RS_ITEMS>(kin, vsrc, kout, vout, n, shift, width, hist, off, partials, s);
            else if constexpr (sizeof(K) == 4)
                radix_pass<K, 10, RS_ITEMS>(kin, vsrc, kout, vout, n, shift, width, hist, off, partials, s);
        }
        shift += width;
        K* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
        flips ^= 1;
    }
    return flips;
}

int radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* tkeys, uint32_t* tvals, int64_t n, int begin_bit,
                     int end_bit, int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s, bool iota) {
    return radix_sort_impl<uint64_t>(keys, vals, tkeys, tvals, n, begin_bit, end_bit, hist, off, partials, iota, false,
                                     s);
}

int radix_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t* tkeys, uint32_t* tvals, int64_t n, int begin_bit,
                       int end_bit, int32_t* hist, int32_t* off, uint32_t* partials, hipStream_t s, bool iota,
                       bool wide) {
    return radix_sort_impl<uint32_t>(keys, vals, tkeys, tvals, n, begin_bit, end_bit, hist, off, partials, iota, wide, s);
}
