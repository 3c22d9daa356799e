// sort.hip — device-wide exclusive scan and stable LSD radix sort (64-bit keys, 32-bit values).
//
// Radix sort: 8-bit digits. Per pass: a histogram kernel (LDS counters per 2048-key tile),
// an exclusive scan of the digit-major histogram, and a scatter kernel that ranks keys
// stably inside each wavefront with ballot-based digit matching (64-lane masks), then
// orders wavefronts of a tile by their position. Deterministic and stable.
#include <algorithm>

#include "lagpm_internal.h"
#include "wave_ops.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

constexpr int RS_THREADS = 256;
// 2048 keys per tile: the scatter's 24 KB of LDS lets 6 tiles share a CU. Measured per 256^3 step
// (six passes): 4096-key tiles 2.41 ms of sort, 3072 2.25, 2048 2.23, 1536 2.30, 1024 2.43
// (shorter digit runs per tile write more partial cache lines).
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-level exclusive scan of SCAN_TILE elements; writes per-block totals, and the grand total
// when this launch is a single tile (the top of the recursion) and `total` is given.
__global__ void __launch_bounds__(SCAN_THREADS) scan_tile_kernel(const uint32_t* __restrict__ in,
                                                                 uint32_t* __restrict__ out, size_t n,
                                                                 uint32_t* __restrict__ block_sums,
                                                                 uint32_t* __restrict__ total) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)tid * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        size_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t incl = wave_incl_scan(s, lane);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t woff = 0, btot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        if (w < wid) woff += wsum[w];
        btot += wsum[w];
    }
    uint32_t run = woff + incl - s;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        size_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (tid == 0 && block_sums) block_sums[blockIdx.x] = btot;
    if (tid == 0 && total && gridDim.x == 1) *total = btot;
}

__global__ void scan_add_kernel(uint32_t* __restrict__ out, size_t n, const uint32_t* __restrict__ block_off) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_off[i / SCAN_TILE];
}

__global__ void scan_total_kernel(const uint32_t* __restrict__ in, const uint32_t* __restrict__ out, size_t n,
                                  uint32_t* __restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = (n ? out[n - 1] + in[n - 1] : 0u);
}

// Radix pass offsets, step 1: one workgroup per digit d scans that digit's column of per-tile
// counts (hist[d * nblocks + t], contiguous in t) into tile offsets within the digit's run, and
// writes the column total to coltot[d]. The scatter adds the digit's base (the totals of the
// smaller digits), so a pass needs one scan launch over its 256 x nblocks counts.
__global__ void __launch_bounds__(SCAN_THREADS) rs_colscan_kernel(const uint32_t* __restrict__ hist,
                                                                  uint32_t* __restrict__ off, uint32_t nblocks,
                                                                  uint32_t* __restrict__ coltot) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const size_t col = (size_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (size_t c0 = 0; c0 < nblocks; c0 += SCAN_TILE) {
        const size_t first = c0 + (size_t)tid * SCAN_ITEMS;
        uint32_t v[SCAN_ITEMS];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            v[k] = first + k < nblocks ? hist[col + first + k] : 0u;
            s += v[k];
        }
        const uint32_t incl = wave_incl_scan(s, lane);
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        uint32_t woff = 0, ctot = 0;
#pragma unroll
        for (int w = 0; w < SCAN_THREADS / 64; w++) {
            if (w < wid) woff += wsum[w];
            ctot += wsum[w];
        }
        uint32_t run = carry + woff + incl - s;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k++) {
            if (first + k < nblocks) off[col + first + k] = run;
            run += v[k];
        }
        carry += ctot;
        __syncthreads();  // wsum is written again by the next chunk
    }
    if (tid == 0) coltot[blockIdx.x] = carry;
}

// ---- radix sort ----
__global__ void __launch_bounds__(RS_THREADS) rs_hist_kernel(const uint64_t* __restrict__ key, size_t n, int shift,
                                                             uint32_t* __restrict__ ghist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int k = 0; k < RS_ITEMS; k++) {
        size_t i = base + (size_t)k * RS_THREADS + tid;
        if (i < n) atomicAdd(&h[(uint32_t)(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    ghist[(size_t)tid * nblocks + blockIdx.x] = h[tid];
}

// Scatter of one tile: keys are ranked stably per wavefront (ballot digit matching),
// placed digit by digit into LDS (tile order: digit, then wave, then position), and written out
// from LDS so that consecutive threads write consecutive slots of one digit's output run
// (coalesced stores; writing from the original positions scatters 64 lanes over 64 runs).
__global__ void __launch_bounds__(RS_THREADS) rs_scatter_kernel(const uint64_t* __restrict__ key,
                                                                const uint32_t* __restrict__ val, size_t n, int shift,
                                                                const uint32_t* __restrict__ goff, uint32_t nblocks,
                                                                const uint32_t* __restrict__ coltot,
                                                                uint64_t* __restrict__ key_out,
                                                                uint32_t* __restrict__ val_out) {
    static_assert(RS_THREADS == 256, "one thread per digit in the offset set-up");
    __shared__ uint32_t dsum[RS_WAVES];
    __shared__ uint32_t wcount[RS_WAVES][256];
    __shared__ uint32_t boff[256];
    __shared__ uint32_t toff[256];
    __shared__ uint32_t wsum[RS_WAVES];
    __shared__ uint64_t skey[RS_TILE];
    __shared__ uint32_t sval[RS_TILE];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // the tile this workgroup scatters: neighbouring tiles write the adjacent ends of each digit's
    // runs, so one XCD takes a contiguous range of tiles and writes those lines whole
    const uint32_t tile = lagpm_wave::xcd_block(blockIdx.x, gridDim.x);
    {
        // this tile's start in digit tid's output run: the digit's base (exclusive prefix of the
        // column totals over the smaller digits) plus the tile's offset within the column
        const uint32_t t = coltot[tid];
        const uint32_t incl = wave_incl_scan(t, lane);
        if (lane == 63) dsum[wid] = incl;
        __syncthreads();
        uint32_t dbase = incl - t;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++)
            if (w < wid) dbase += dsum[w];
        boff[tid] = dbase + goff[(size_t)tid * nblocks + tile];
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) wcount[w][tid] = 0;
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t tile0 = (size_t)tile * RS_TILE;
    const size_t seg = tile0 + (size_t)wid * (RS_ITEMS * 64);
    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        size_t i = seg + (size_t)it * 64 + lane;
        bool act = i < n;
        k[it] = act ? key[i] : 0ull;
        v[it] = act ? val[i] : 0u;
        uint32_t d = (uint32_t)(k[it] >> shift) & 255u;
        uint64_t m = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t r = (uint32_t)__popcll(m & lt);
        uint32_t c = (uint32_t)__popcll(m);
        uint32_t before = act ? wcount[wid][d] : 0u;
        rank[it] = act ? before + r : 0xffffffffu;
        __builtin_amdgcn_wave_barrier();
        if (act && r == 0) wcount[wid][d] = before + c;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: the waves' exclusive prefix (in place) and the tile total, then the tile's
    // exclusive scan over digits (RS_THREADS == 256: thread d owns digit d)
    {
        const int d = tid;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            uint32_t t = wcount[w][d];
            wcount[w][d] = run;
            run += t;
        }
        int wt;
        const uint32_t ex = (uint32_t)lagpm_wave::wave_excl_scan((int)run, lane, wt);
        if (lane == 0) wsum[wid] = (uint32_t)wt;
        __syncthreads();
        uint32_t wo = 0;
        for (int w = 0; w < wid; w++) wo += wsum[w];
        toff[d] = wo + ex;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < RS_ITEMS; it++) {
        if (rank[it] == 0xffffffffu) continue;
        const uint32_t d = (uint32_t)(k[it] >> shift) & 255u;
        const uint32_t lp = toff[d] + wcount[wid][d] + rank[it];
        skey[lp] = k[it];
        sval[lp] = v[it];
    }
    __syncthreads();
    const uint32_t tn = (uint32_t)(n - tile0 < (size_t)RS_TILE ? n - tile0 : (size_t)RS_TILE);
    for (uint32_t j = tid; j < tn; j += RS_THREADS) {
        const uint64_t kk = skey[j];
        const uint32_t d = (uint32_t)(kk >> shift) & 255u;
        const size_t dst = (size_t)boff[d] + (j - toff[d]);
        key_out[dst] = kk;
        val_out[dst] = sval[j];
    }
}

// recursive exclusive scan using tmp (enough room for all levels); the grand total into *total
// (when given) from the single-tile scan at the top of the recursion
int scan_rec(lagpm_ctx* ctx, hipStream_t st, const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp,
             uint32_t* total = nullptr) {
    size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 0) return LAGPM_OK;
    uint32_t* sums = tmp;
    uint32_t* sums_scan = tmp + nb;
    hipLaunchKernelGGL(scan_tile_kernel, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, st, in, out, n,
                       nb > 1 ? sums : (uint32_t*)nullptr, nb > 1 ? (uint32_t*)nullptr : total);
    HIPCHK(ctx, hipGetLastError());
    if (nb > 1) {
        RCHK(scan_rec(ctx, st, sums, sums_scan, nb, tmp + 2 * nb, total));
        hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n,
                           sums_scan);
        HIPCHK(ctx, hipGetLastError());
    }
    return LAGPM_OK;
}

size_t scan_tmp_elems(size_t n) {
    size_t t = 1;
    while (n > (size_t)SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        t += 2 * n;
    }
    return t + 2 * ((n + SCAN_TILE - 1) / SCAN_TILE) + 64;
}

}  // namespace

size_t lagpm_scan_tmp_bytes(size_t n) { return (scan_tmp_elems(n) + 4) * sizeof(uint32_t); }

size_t lagpm_sort_tmp_bytes(size_t n) {
    size_t nblocks = (n + RS_TILE - 1) / RS_TILE;
    size_t hist_elems = (size_t)256 * std::max<size_t>(nblocks, 1);
    return (2 * hist_elems + 256) * sizeof(uint32_t);  // counts, offsets, column totals
}

int lagpm_exclusive_scan_u32(lagpm_ctx* ctx, hipStream_t st, uint32_t* tmp, const uint32_t* in, uint32_t* out, size_t n,
                             uint32_t* total_dev) {
    if (n > 0) RCHK(scan_rec(ctx, st, in, out, n, tmp + 4, total_dev));
    if (total_dev && n == 0) {
        hipLaunchKernelGGL(scan_total_kernel, dim3(1), dim3(64), 0, st, in, out, n, total_dev);
        HIPCHK(ctx, hipGetLastError());
    }
    return LAGPM_OK;
}

// Stable sort of (key, val) pairs on `bits` bits of key from bit lo_bit up (rounded up to 8).
int lagpm_radix_sort_pairs(lagpm_ctx* ctx, hipStream_t st, uint32_t* base, uint64_t* key, uint64_t* key_alt, uint32_t* val,
                           uint32_t* val_alt, size_t n, int bits, bool* result_in_alt, int lo_bit) {
    *result_in_alt = false;
    if (n <= 1) return LAGPM_OK;
    uint32_t nblocks = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    size_t hist_elems = (size_t)256 * nblocks;
    uint32_t* ghist = base;               // per-tile digit counts, digit-major
    uint32_t* goff = ghist + hist_elems;  // tile offsets within each digit's run
    uint32_t* coltot = goff + hist_elems;  // per-digit totals
    uint64_t *ki = key, *ko = key_alt;
    uint32_t *vi = val, *vo = val_alt;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; p++) {
        int shift = lo_bit + 8 * p;
        hipLaunchKernelGGL(rs_hist_kernel, dim3(nblocks), dim3(RS_THREADS), 0, st, ki, n, shift, ghist, nblocks);
        HIPCHK(ctx, hipGetLastError());
        hipLaunchKernelGGL(rs_colscan_kernel, dim3(256), dim3(SCAN_THREADS), 0, st, ghist, goff, nblocks, coltot);
        HIPCHK(ctx, hipGetLastError());
        hipLaunchKernelGGL(rs_scatter_kernel, dim3(nblocks), dim3(RS_THREADS), 0, st, ki, vi, n, shift, goff,
                           nblocks, coltot, ko, vo);
        HIPCHK(ctx, hipGetLastError());
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *result_in_alt = (ki != key);
    return LAGPM_OK;
}
