// bins.hip -- points binned by brick cell with an LSD radix sort that moves
// the point records themselves.
//
// A point's record is (x, y, z, index) after the translation, 16 bytes; an
// invalid point's is (-1, -1, -1, index).  Its key is its cell in the bin
// grid (the leaf grid grown by one cell), or the grid's cell count for
// invalid points and points outside the grid, so those sort last.  The key
// is recomputed from the record in every pass instead of being stored, and
// the first pass reads the caller's points in index order, so no pass reads
// anything at random: per point 12 B (first pass, twice: histogram and
// scatter) or 16 B read and 16 B written per pass.  Sorting (key, index)
// pairs and then gathering the points by index instead reads each 12-byte
// point from a random place, a whole cache line for 12 bytes.
//
// Per pass (digits of at most 10 bits, as few passes as that allows, of
// near-equal width): a per-tile digit histogram, an exclusive scan over the
// digit-major histogram, and a scatter in which each wave ranks a quarter of
// its tile (64 records per round, a ballot per digit bit, its own running
// digit counts in LDS), the tile is reordered by digit in LDS and every
// digit's run is written to its place.  Stable: every bin ends in ascending
// point index.  The last pass also writes the sorted keys (for the bin
// offsets) and, when asked, each cell's extra-brick reach (the reach of the
// points of one cell, OR-ed per wave, one atomicOr per run of a cell).
#include "gvdbx_kernels.h"
#include "gvdbx_reach.h"

namespace {

constexpr int BS_THREADS = 256;
constexpr int BS_ITEMS = 16;
constexpr int BS_TILE = BS_THREADS * BS_ITEMS;
constexpr int BS_WAVES = BS_THREADS / 64;

struct BinSortIn {
    PointSrc src;          // first pass: the caller's points
    float tx, ty, tz;
    const float4* rec;     // later passes: the previous pass's records
    CellGrid gb;
    int logr0;
    uint32_t none;         // gb.count(): key of invalid and outside points
};

__device__ __forceinline__ uint32_t rec_key(const float4& r, const BinSortIn& in) {
    if (!(r.x >= 0.0f)) return in.none;
    const int cx = gvdbx_cell(r.x, in.logr0), cy = gvdbx_cell(r.y, in.logr0), cz = gvdbx_cell(r.z, in.logr0);
    if (!in.gb.sparse()) {
        // a dense grid holds at most 2^28 cells: 32-bit index arithmetic is
        // exact, and one unsigned compare per axis covers both box faces
        const uint32_t ux = (uint32_t)cx - (uint32_t)in.gb.ox, uy = (uint32_t)cy - (uint32_t)in.gb.oy,
                       uz = (uint32_t)cz - (uint32_t)in.gb.oz;
        if (ux >= (uint32_t)in.gb.nx || uy >= (uint32_t)in.gb.ny || uz >= (uint32_t)in.gb.nz) return in.none;
        return (ux * (uint32_t)in.gb.ny + uy) * (uint32_t)in.gb.nz + uz;
    }
    const int64_t b = in.gb.find(cx, cy, cz);
    return b >= 0 ? (uint32_t)b : in.none;
}

template <bool FIRST>
__device__ __forceinline__ float4 load_rec(const BinSortIn& in, int64_t i) {
    if constexpr (FIRST) {
        float w[3];
        if (gvdbx_point(in.src, i, in.tx, in.ty, in.tz, w[0], w[1], w[2]))
            return make_float4(w[0], w[1], w[2], __int_as_float((int)i));
        return make_float4(-1.0f, -1.0f, -1.0f, __int_as_float((int)i));
    } else {
        return in.rec[i];
    }
}

template <int DB, bool FIRST>
__global__ __launch_bounds__(BS_THREADS) void k_bs_hist(BinSortIn in, int64_t n, int shift, uint32_t mask,
                                                         int64_t ntiles, int32_t* hist) {
    constexpr int B = 1 << DB;
    __shared__ int32_t cnt[B];
    for (int b = threadIdx.x; b < B; b += BS_THREADS) cnt[b] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * BS_TILE;
    float4 r[BS_ITEMS];
#pragma unroll
    for (int u = 0; u < BS_ITEMS; u++) {   // all loads first, then the LDS counts
        const int64_t i = t0 + u * BS_THREADS + threadIdx.x;
        if (i < n) r[u] = load_rec<FIRST>(in, i);
    }
#pragma unroll
    for (int u = 0; u < BS_ITEMS; u++)
        if (t0 + u * BS_THREADS + threadIdx.x < n) atomicAdd(&cnt[(rec_key(r[u], in) >> shift) & mask], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += BS_THREADS) hist[(int64_t)b * ntiles + blockIdx.x] = cnt[b];
}

template <int DB, bool FIRST, bool LAST>
__global__ __launch_bounds__(BS_THREADS) void k_bs_scatter(BinSortIn in, int64_t n, int shift, uint32_t mask,
                                                            int64_t ntiles, const int32_t* hist, const int32_t* off,
                                                            float4* out, uint32_t* okeys, ReachParams rp,
                                                            uint32_t* reach0) {
    constexpr int B = 1 << DB;
    constexpr int R = B / BS_THREADS > 0 ? B / BS_THREADS : 1;
    static_assert(B >= BS_THREADS, "one bin or more per thread");
    __shared__ int32_t gbase[B];             // global output start of each digit for this tile
    __shared__ int32_t tstart[B];            // start of each digit inside the tile
    __shared__ int32_t wcnt[BS_WAVES][B];    // per wave: running digit counts, then offsets
    __shared__ int32_t wtot[BS_WAVES];
    __shared__ float4 lrec[BS_TILE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * BS_TILE;
    const int tn = (int)min((int64_t)BS_TILE, n - t0);
    int c[R];
    int csum = 0;
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int b = tid * R + j;
        gbase[b] = off[(int64_t)b * ntiles + blockIdx.x];
        for (int w = 0; w < BS_WAVES; w++) wcnt[w][b] = 0;
        c[j] = hist[(int64_t)b * ntiles + blockIdx.x];
        csum += c[j];
    }
    {   // exclusive scan of this tile's digit counts
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
    // wave w ranks the tile's records [w * 64 * ITEMS, (w + 1) * 64 * ITEMS)
    // against its own running digit counts (no block barrier between rounds)
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int wbase = wave * 64 * BS_ITEMS;
    float4 rr[BS_ITEMS];
    int rk[BS_ITEMS];
    int dg[BS_ITEMS];
#pragma unroll
    for (int u = 0; u < BS_ITEMS; u++) {   // all loads in flight at once
        const int li = wbase + u * 64 + lane;
        if (li < tn) rr[u] = load_rec<FIRST>(in, t0 + li);
    }
    int32_t* wc = wcnt[wave];
#pragma unroll
    for (int u = 0; u < BS_ITEMS; u++) {
        const bool ok = wbase + u * 64 + lane < tn;
        const int d = ok ? (int)((rec_key(rr[u], in) >> shift) & mask) : 0;
        dg[u] = d;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            const uint64_t m = __ballot(ok && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        const int before = ok ? wc[d] : 0;
        rk[u] = before + rank;
        if (ok && rank == 0) wc[d] = before + __popcll(peers);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int b = tid * R + j;
        int acc = 0;
        for (int w = 0; w < BS_WAVES; w++) {
            const int t = wcnt[w][b];
            wcnt[w][b] = acc;
            acc += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BS_ITEMS; u++)
        if (wbase + u * 64 + lane < tn) lrec[tstart[dg[u]] + wc[dg[u]] + rk[u]] = rr[u];
    __syncthreads();
    // tile order is digit order: consecutive threads write consecutive places
    // of one digit's run (after the last pass the tile is in full key order)
    for (int j0 = 0; j0 < tn; j0 += BS_THREADS) {   // wave-uniform steps (the reach shuffles)
        const int j = j0 + tid;
        uint32_t key = in.none;
        float4 r = make_float4(-1.0f, -1.0f, -1.0f, 0.0f);
        if (j < tn) {
            r = lrec[j];
            key = rec_key(r, in);
            const int d = (int)((key >> shift) & mask);
            const int64_t pos = (int64_t)gbase[d] + (j - tstart[d]);
            out[pos] = r;
            if (LAST) okeys[pos] = key;
        }
        if (LAST && reach0) {
            int rk2 = -1;
            uint32_t m = 0;
            if (key != in.none) {
                const float w[3] = {r.x, r.y, r.z};
                reach_point(w, __float_as_int(r.w), rp, rk2, m);
            }
            reach_commit(rk2, m, reach0);
        }
    }
}

template <int DB>
void bs_pass(const BinSortIn& in, bool first, bool last, int64_t n, int shift, int width, int32_t* hist,
             int32_t* off, uint32_t* partials, float4* out, uint32_t* keys, const ReachParams& rp, uint32_t* reach0,
             hipStream_t s) {
    const int64_t ntiles = std::max<int64_t>((n + BS_TILE - 1) / BS_TILE, 1);
    const int64_t nh = ntiles << DB;
    const int64_t stiles = scan_tiles(nh + 1);
    const uint32_t mask = (1u << width) - 1u;
    const dim3 g((unsigned)ntiles), b(BS_THREADS);
    if (first)
        hipLaunchKernelGGL((k_bs_hist<DB, true>), g, b, 0, s, in, n, shift, mask, ntiles, hist);
    else
        hipLaunchKernelGGL((k_bs_hist<DB, false>), g, b, 0, s, in, n, shift, mask, ntiles, hist);
    launch_count_scan_reduce(hist, nh, partials, partials + stiles, s);
    launch_count_scan_apply(hist, nh, partials, off, s);
    if (first && last)
        hipLaunchKernelGGL((k_bs_scatter<DB, true, true>), g, b, 0, s, in, n, shift, mask, ntiles, hist, off, out, keys,
                           rp, reach0);
    else if (first)
        hipLaunchKernelGGL((k_bs_scatter<DB, true, false>), g, b, 0, s, in, n, shift, mask, ntiles, hist, off, out,
                           keys, rp, reach0);
    else if (last)
        hipLaunchKernelGGL((k_bs_scatter<DB, false, true>), g, b, 0, s, in, n, shift, mask, ntiles, hist, off, out,
                           keys, rp, reach0);
    else
        hipLaunchKernelGGL((k_bs_scatter<DB, false, false>), g, b, 0, s, in, n, shift, mask, ntiles, hist, off, out,
                           keys, rp, reach0);
}

}  // namespace

int bin_sort_records(PointSrc src, int64_t n, float tx, float ty, float tz, CellGrid gb, int logr0, float4* buf0,
                     float4* buf1, uint32_t* keys, int32_t* hist, int32_t* off, uint32_t* partials,
                     const ReachParams* rp, uint32_t* reach0, hipStream_t s) {
    if (n <= 0) return 0;
    BinSortIn in{};
    in.src = src;
    in.tx = tx, in.ty = ty, in.tz = tz;
    in.gb = gb;
    in.logr0 = logr0;
    in.none = (uint32_t)gb.count();
    int bits = 1;
    while (((int64_t)1 << bits) <= (int64_t)in.none) bits++;   // keys 0 .. none
    const int np = (bits + 9) / 10;
    const ReachParams r = rp ? *rp : ReachParams{};
    float4* bufs[2] = {buf0, buf1};
    int cur = 0;
    for (int p = 0, shift = 0; p < np; p++) {
        const int width = bits / np + (p < bits % np ? 1 : 0);
        const bool first = p == 0, last = p == np - 1;
        in.rec = first ? nullptr : bufs[cur ^ 1];
        float4* out = bufs[cur];
        uint32_t* rch = last && rp ? reach0 : nullptr;
        if (width <= 8)
            bs_pass<8>(in, first, last, n, shift, width, hist, off, partials, out, keys, r, rch, s);
        else if (width == 9)
            bs_pass<9>(in, first, last, n, shift, width, hist, off, partials, out, keys, r, rch, s);
        else
            bs_pass<10>(in, first, last, n, shift, width, hist, off, partials, out, keys, r, rch, s);
        shift += width;
        cur ^= 1;
    }
    return cur ^ 1;   // the buffer the last pass wrote
}
