// tree.hip — short-range tree correction: Morton sort, linear octree, node moments, target
// groups and the wavefront-cooperative walk with the tabulated split force law.
//
// Reference semantics (paths relative to the reference tree):
//   driver       Cosmos/treecorrection.c:225-475 (chaining mesh of width >= cutoff, 27 neighbours)
//   tree         OST/ost.ver2.c:36-196 (monopole + quadrupole about the COM, rmax = max |x - COM|)
//   opening      OST/ost.ver2.c:296-306 (open when d^2 < rmax^2 / theta^2)
//   node force   OST/ost.ver2.c:247-277 (AddNode, with the y/z components of defect D1 restored)
//   pair force   OST/ost.ver2.c:279-293 (AddPar); both use the cutoff test d^2 <= rsphere^2
//   force law    OST/Model4TreeForce.c:34-46 (GetPointGrav / GetCellGrav, 4096-entry linear tables)
//
// MI355X design (not the reference's): one global octree per box instead of one tree per
// 4x4x4 cell. Particles are sorted by a 48-bit Morton key; the chaining cells are the
// octree nodes of level L (2^L per axis, width n/2^L >= cutoff). A node is the run of
// particles sharing a key prefix; nodes are numbered in pre-order by (first particle, level),
// which a per-particle count + exclusive scan produces without any level-by-level pass.
// Target groups are subtrees of <= 128 particles; one wavefront walks the 27 neighbouring
// cell trees for its group with a group-conservative opening test, stages accepted nodes
// and opened leaves in LDS, and evaluates the interactions lane-per-target.
#include "lagpm_internal.h"
#include "wave_ops.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {
using namespace lagpm_wave;

#ifndef WALK_THREADS_PER_CU
#define WALK_THREADS_PER_CU 1024
#endif
constexpr int WALK_THREADS = WALK_THREADS_PER_CU;  // one workgroup per CU (LDS: shared table + 16 wave lists)
constexpr int WALK_WAVES = WALK_THREADS / 64;
#ifndef WALK_PP_TAKE
#define WALK_PP_TAKE 4
#endif
#ifndef WALK_LSTACK
#define WALK_LSTACK 256
#endif
constexpr int PP_TAKE = WALK_PP_TAKE;  // particles a lane appends per round
constexpr int PP_CAP = 64 * PP_TAKE + 64;  // one round + 63 left over
constexpr int PC_CAP = 128;
constexpr int STACK_MAX = 8192;  // per-wave walk stack entries
constexpr int WALK_XCDS = LAGPM_WALK_XCDS;
constexpr int LSTACK = WALK_LSTACK;  // the first LSTACK entries of a wave's stack live in LDS, the rest in global memory
constexpr int WALK_MAX_ITERS = 1 << 20;  // per-group bound on 64-node pop rounds
// rmax: a leaf and an inner node of up to RMAX_EXACT_MAX particles get the exact largest distance
// of their particles from the com (the reference's node size); larger nodes get the bound
// max(|com_child - com| + rmax_child), which only opens such nodes somewhat earlier.
#ifndef RMAX_EXACT_MAX
#define RMAX_EXACT_MAX 128
#endif
constexpr float RAN_F = (float)(6.0 / 4096.0);
constexpr int TAB_LDS = LAGPM_LOG_BINS > 4096 ? LAGPM_LOG_BINS : 4096;

// Table bin and offset of a pair at squared distance r2 (see law_tables): uniform r bins need
// the square root; log r^2 bins take the exponent and 8 mantissa bits of r2.
template <bool LOGT>
__device__ inline int table_bin(float r2, float& t, float lam) {
    if constexpr (LOGT) {
        // Each bin holds the line c + s * r2 (bin 0 runs from r2 = 0, so it serves every r2 below
        // its node); bins beyond the cutoff hold zeros, so far pairs and the far-away dummies of a
        // partial chunk add nothing without a compare. r2 above the table is clamped to its last bin.
        // lam r2 >= 0, so its bit pattern orders like the value: one unsigned clamp of the bits to
        // [bin 0's node, just below 64] (a v_med3_u32) replaces a float clamp and an integer max.
        constexpr uint32_t LO = (uint32_t)LAGPM_LOG_BASE << LAGPM_LOG_SHIFT;
        constexpr uint32_t HI = ((uint32_t)(LAGPM_LOG_BASE + LAGPM_LOG_BINS) << LAGPM_LOG_SHIFT) - 1u;
        const uint32_t b = min(max(__float_as_uint(lam * r2), LO), HI);
        t = r2;
        // The bin, bits 15..30 of b, as one bit-field extract: the table address is then one
        // shift-and-add from it. Written as (b >> 15) the compiler turns the shift and the
        // address's scaling by 8 into a shift, a mask and an add: one more operation per pair.
        static_assert(LAGPM_LOG_SHIFT == 15, "the extract below takes bits 15..30");
        uint32_t f;
        asm("v_bfe_u32 %0, %1, 15, 16" : "=v"(f) : "v"(b));
        return (int)f - (int)LAGPM_LOG_BASE;
    } else {
        const float r = __builtin_amdgcn_sqrtf(r2);
        const int nt = min((int)(r * (1.0f / RAN_F)), 4095);
        t = r - (float)nt * RAN_F;
        return nt;
    }
}
template <bool LOGT>
__device__ inline int table_bin64(double r2, double& t, float lam) {
    if constexpr (LOGT) {
        const double r2c = fmin((double)lam * r2, 63.99);
        t = r2;  // bins hold c + s * r2, zero beyond the cutoff (see table_bin)
        return max((int)(__float_as_uint((float)r2c) >> LAGPM_LOG_SHIFT) - (int)LAGPM_LOG_BASE, 0);
    } else {
        const double ran = 6.0 / 4096.0;
        const double r = sqrt(r2);
        const int nt = min((int)(r / ran), 4095);
        t = r - nt * ran;
        return nt;
    }
}

__device__ inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0xffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
__device__ inline uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x | (x >> 4)) & 0x100f00f00f00f00full;
    x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x | (x >> 16)) & 0x1f00000000ffffull;
    x = (x | (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}
__device__ inline uint32_t quant(double x, double s) {
    double q = floor(x * s);
    if (q < 0) q = 0;
    if (q > 65535.0) q = 65535.0;
    return (uint32_t)q;
}
__device__ inline int key_shift(int level) { return 3 * (LAGPM_KEY_BITS - level); }
// number of leading octal digits two keys share (16 when equal)
__device__ inline int common_digits(uint64_t a, uint64_t b) {
    uint64_t d = a ^ b;
    if (d == 0) return LAGPM_KEY_BITS;
    int p = 63 - __clzll(d);
    return (47 - p) / 3;
}
// first index in [lo, hi) whose prefix (key >> sh) exceeds pfx
__device__ inline size_t upper_bound_pfx(const uint64_t* __restrict__ K, size_t lo, size_t hi, uint64_t pfx, int sh) {
    while (lo < hi) {
        size_t mid = lo + ((hi - lo) >> 1);
        if ((K[mid] >> sh) > pfx) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}
// first index in [lo, hi) whose prefix is >= pfx
__device__ inline size_t lower_bound_pfx(const uint64_t* __restrict__ K, size_t lo, size_t hi, uint64_t pfx, int sh) {
    while (lo < hi) {
        size_t mid = lo + ((hi - lo) >> 1);
        if ((K[mid] >> sh) >= pfx) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Morton keys of the positions and the identity permutation. (The Hubble drag of the step's front,
// v *= drag[0], Cosmos/globalT.c:102-110, rides in the gather after the sort.)
__global__ void keys_kernel(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                            size_t np, double s, uint64_t* __restrict__ key, uint32_t* __restrict__ perm) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    key[i] = spread3(quant(x[i], s)) | (spread3(quant(y[i], s)) << 1) | (spread3(quant(z[i], s)) << 2);
    perm[i] = (uint32_t)i;
}

// all particle arrays into the new order in one pass (one permutation load, seven gathers in flight)
struct gather_set {
    const double* src[6];
    double* dst[6];
    const uint32_t* id;
    uint32_t* id_dst;
};
// The sources of neighbouring outputs share cache lines (the order changes little between steps),
// so the gather gives each XCD one contiguous range of outputs (lagpm_wave::xcd_block): the lines
// those outputs read are fetched into one L2 instead of into several. With drag, the velocities
// are multiplied by drag[0] on the way (the Hubble drag of the step's front), a pass over them
// that the keys would otherwise make.
__global__ void gather_all_kernel(gather_set g, const uint32_t* __restrict__ perm, size_t np,
                                  const double* __restrict__ drag) {
    const size_t i = (size_t)lagpm_wave::xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t p = perm[i];
    double v[6];
#pragma unroll
    for (int a = 0; a < 6; a++) v[a] = g.src[a][p];
    const uint32_t id = g.id[p];
    if (drag) {
        const double f = drag[0];
        v[3] *= f;
        v[4] *= f;
        v[5] *= f;
    }
#pragma unroll
    for (int a = 0; a < 6; a++) g.dst[a][i] = v[a];
    g.id_dst[i] = id;
}

// The same gather in two halves, for a step front that runs them on two streams: the positions
// (which the deposit and the tree build wait for) and the velocities with the ids and the drag
// (which only the PM kick and the drift need, so they load beside the deposit and the build).
__global__ void gather_pos_kernel(gather_set g, const uint32_t* __restrict__ perm, size_t np) {
    const size_t i = (size_t)lagpm_wave::xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t p = perm[i];
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = g.src[a][p];
#pragma unroll
    for (int a = 0; a < 3; a++) g.dst[a][i] = v[a];
}
__global__ void gather_vel_kernel(gather_set g, const uint32_t* __restrict__ perm, size_t np,
                                  const double* __restrict__ drag) {
    const size_t i = (size_t)lagpm_wave::xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t p = perm[i];
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = g.src[3 + a][p];
    const uint32_t id = g.id[p];
    if (drag) {
        const double f = drag[0];
#pragma unroll
        for (int a = 0; a < 3; a++) v[a] *= f;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) g.dst[3 + a][i] = v[a];
    g.id_dst[i] = id;
}

__device__ inline void tree_reset(int t, uint32_t* __restrict__ dsz, unsigned long long* __restrict__ counters) {
    if (t == DSZ_NODES || t == DSZ_GROUPS || t == DSZ_OVERFLOW || t == DSZ_DEEPER || t == DSZ_WALKOVF) dsz[t] = 0u;
    if ((t >= DSZ_LVL && t < DSZ_LOFF) || (t >= DSZ_LCUR && t < DSZ_COUNT)) dsz[t] = 0u;
    if (t < 4) counters[t] = 0ull;
}

// Everything the build needs cleared or packed before the level-L cells, in one pass: the build's
// sizes and counters, per-particle node counts, empty cell roots, and (FP32 tree, pp non-null)
// the pp copy of the positions in units of the particle mass.
// pp is the FP32 copy of the positions for the tree, scaled by sc (rounded once from the FP64
// product): with the log-r^2 force tables sc = sqrt(lam), so a pair's r^2 in these units is lam r^2,
// the table variable, and needs no multiply in the pair loop. Everything the FP32 tree derives
// from pp (centres, second moments, node sizes, group boxes) is in the same units.
__global__ void tree_prep_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, size_t np, double sc, float4* __restrict__ pp,
                                 uint32_t* __restrict__ ncount, int* __restrict__ cellroot, int* __restrict__ wcellroot,
                                 size_t ncell, uint32_t* __restrict__ dsz, unsigned long long* __restrict__ counters) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 64) tree_reset(threadIdx.x, dsz, counters);
    if (i < np) {
        if (pp) pp[i] = make_float4((float)(x[i] * sc), (float)(y[i] * sc), (float)(z[i] * sc), 1.0f);
        ncount[i] = 0u;
    }
    if (i < ncell) {
        cellroot[i] = -1;
        if (wcellroot) wcellroot[i] = -1;
    }
}

// node count on the device, clamped to the array capacity
__device__ inline size_t dev_count(const uint32_t* __restrict__ dsz, int which, size_t cap) {
    size_t c = dsz[which];
    return c < cap ? c : cap;
}

// ---------------------------------------------------------------------------------------
// Top-down build, one level per launch. A node record is (first particle, count, level, children)
// in the order it was made; the records of level l are contiguous (dsz[DSZ_LOFF + l], count
// dsz[DSZ_LVL + l]), which is also the level order the moments use. Cells of the chaining level L
// are the roots; a node of more than nleaf particles above level 16 splits into its non-empty
// octants, found by 8 lanes bisecting the node's key range in parallel, one octant boundary each.
// Afterwards every node gets its pre-order index (first particle, then level): nodes that begin at
// particle f are a chain of consecutive levels from lmin(f) = max(L, 1 + digits f shares with f-1),
// so index = (nodes beginning before f) + level - lmin(f), with the first term from a scan of
// per-particle node counts. The index does not depend on the order in which records were made,
// so the numbering (and everything downstream) is deterministic.

// Append for the lanes whose flag is set: one atomic per wave, slots in lane order. Every lane of
// the wave must call it.
__device__ inline uint32_t wave_append(bool make, uint32_t* counter) {
    const unsigned long long m = __ballot(make);
    if (!m) return 0u;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0u;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    return base + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
// the level-L cells that hold particles, one thread per cell (Morton index c = the key prefix)
__global__ void cells_kernel(const uint64_t* __restrict__ K, size_t N, int L, size_t cap, uint32_t* __restrict__ dsz,
                             int4* __restrict__ rec, uint32_t* __restrict__ ncount) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ncell = (size_t)1 << (3 * L);
    uint32_t a = 0u, cnt = 0u;
    if (c < ncell) {
        const int sh = key_shift(L);
        a = (uint32_t)lower_bound_pfx(K, 0, N, (uint64_t)c, sh);
        cnt = (uint32_t)(upper_bound_pfx(K, a, N, (uint64_t)c, sh) - a);
    }
    const bool make = cnt > 0u;
    const uint32_t r = wave_append(make, &dsz[DSZ_NODES]);
    if (make) {
        if (r < cap) {
            rec[r] = make_int4((int)a, (int)cnt, L, 0);
            ncount[a] += 1u;  // one node per first particle per level: no race
        } else {
            dsz[DSZ_OVERFLOW] = 1u;
        }
    }
}
// level l's records: from the previous level's end to the records made so far (clamped to cap)
__global__ void level_mark_kernel(uint32_t* __restrict__ dsz, size_t cap, int l, int first_level) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t made = dsz[DSZ_NODES] < cap ? dsz[DSZ_NODES] : (uint32_t)cap;
    const uint32_t s = l == first_level ? 0u : dsz[DSZ_LOFF + l - 1] + dsz[DSZ_LVL + l - 1];
    dsz[DSZ_LOFF + l] = s;
    dsz[DSZ_LVL + l] = made - s;
}
// split the level-l records: 8 lanes per record, lane j finds where octant j begins. New
// records are appended with one global atomic per workgroup and pass (a single counter that
// every wave hit directly serialised ~10^5 atomics per level).
constexpr int SPLIT_THREADS = 256;
// At the depth limit (l == lmax) nothing is split: a record that would be is flagged instead.
__global__ void __launch_bounds__(SPLIT_THREADS) split_kernel(const uint64_t* __restrict__ K, int l, int lmax, int nleaf,
                                                              size_t cap, uint32_t* __restrict__ dsz, int4* __restrict__ rec,
                                                              int* __restrict__ rch0, int* __restrict__ rch1,
                                                              uint32_t* __restrict__ ncount, int* __restrict__ rfc) {
    __shared__ uint32_t wcnt[SPLIT_THREADS / 64];
    __shared__ uint32_t bbase;
    const uint32_t n = dsz[DSZ_LVL + l], o = dsz[DSZ_LOFF + l];
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = (int)(threadIdx.x & 7), lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
    const size_t stride = (size_t)gridDim.x * blockDim.x / 8;
    const size_t block_first = (size_t)blockIdx.x * blockDim.x / 8;
    // every thread of the workgroup runs the same passes (the appends below are block-wide)
    for (size_t pass = 0; block_first + pass < n; pass += stride) {
        const size_t r = tid / 8 + pass;
        const bool valid = r < n;
        const int4 nd = valid ? rec[o + r] : make_int4(0, 0, 0, 0);
        const bool need = valid && nd.y > nleaf && l < LAGPM_KEY_BITS;
        if (need && l >= lmax) {
            dsz[DSZ_DEEPER] = 1u;
            dsz[DSZ_OVERFLOW] = 1u;  // the walk and the drift skip this box until it is redone
        }
        const bool split = need && l < lmax;
        uint32_t b = 0u;
        if (split) {
            const uint64_t pfx = ((K[nd.x] >> key_shift(l)) << 3) | (uint64_t)sub;
            b = (uint32_t)lower_bound_pfx(K, (size_t)nd.x, (size_t)nd.x + nd.y, pfx, key_shift(l + 1));
        }
        const uint32_t bn = (uint32_t)__shfl_down((int)b, 1, 64);
        const uint32_t e = sub == 7 ? (uint32_t)(nd.x + nd.y) : bn;
        const bool make = split && e > b;
        const unsigned long long wm = __ballot(make);
        if (lane == 0) wcnt[wid] = (uint32_t)__popcll(wm);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < SPLIT_THREADS / 64; w++) t += wcnt[w];
            bbase = t ? atomicAdd(&dsz[DSZ_NODES], t) : 0u;
        }
        __syncthreads();
        uint32_t rr = bbase;
        for (int w = 0; w < wid; w++) rr += wcnt[w];
        rr += __builtin_amdgcn_mbcnt_hi((unsigned)(wm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)wm, 0u));
        __syncthreads();  // wcnt and bbase are rewritten by the next pass
        if (make) {
            if (rr < cap) {
                rec[rr] = make_int4((int)b, (int)(e - b), l + 1, 0);
                ncount[b] += 1u;
            } else {
                dsz[DSZ_OVERFLOW] = 1u;
            }
        }
        // the record's children, in octant order: lane j's rank among its group's non-empty octants
        const unsigned long long mm = __ballot(make);
        const uint32_t grp = (uint32_t)(mm >> (threadIdx.x & 56)) & 0xffu;
        const int ci = __popc(grp & ((1u << sub) - 1u));
        const size_t R = (size_t)o + r;
        if (make) {
            if (ci < 4) rch0[4 * R + ci] = (int)b;
            else rch1[4 * R + ci - 4] = (int)b;
            // the record's children are consecutive records (a record's lanes append in octant
            // order, records in lane order): the first one's index is all the walk needs
            if (ci == 0 && rfc) rfc[R] = (int)rr;
        }
        if (valid && sub == 0) rec[R].w = __popc(grp);
    }
}
__device__ inline int chain_first_level(const uint64_t* __restrict__ K, uint32_t f, int L) {
    const int cpl = f == 0 ? -1 : common_digits(K[f - 1], K[f]);
    return max(L, cpl + 1);
}
// pre-order numbering: info, child and parent links, cell roots and the level order
__global__ void number_kernel(const uint64_t* __restrict__ K, int L, size_t cap, uint32_t* __restrict__ dsz,
                              const int4* __restrict__ rec, const int* __restrict__ rch0, const int* __restrict__ rch1,
                              const uint32_t* __restrict__ noff, int4* __restrict__ info, int* __restrict__ child,
                              int* __restrict__ parent, int* __restrict__ cellroot, int mcell, uint32_t* __restrict__ order,
                              const int* __restrict__ rfc, int4* __restrict__ wlink, int* __restrict__ wcellroot) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dsz[DSZ_OVERFLOW] || r >= dev_count(dsz, DSZ_NODES, cap)) return;  // an overflowed build is redone
    const int4 nd = rec[r];  // (first, count, level, children)
    const uint32_t v = noff[nd.x] + (uint32_t)(nd.z - chain_first_level(K, (uint32_t)nd.x, L));
    if ((size_t)v >= cap) return;  // cannot happen for sorted keys (the chain levels are consecutive)
    order[r] = v;
    info[v] = make_int4(nd.x, nd.y, nd.w, nd.z);
    // the FP32 walk's links, by record: first particle, count, first child record, children
    if (wlink) wlink[r] = make_int4(nd.x, nd.y, nd.w > 0 ? rfc[r] : 0, nd.w);
    if (nd.z == L) {
        parent[v] = -1;
        const uint64_t cell = K[nd.x] >> key_shift(L);
        const size_t ci = compact3(cell) + (size_t)mcell * (compact3(cell >> 1) + (size_t)mcell * compact3(cell >> 2));
        cellroot[ci] = (int)v;
        if (wcellroot) wcellroot[ci] = (int)r;
    }
    const uint32_t nodes = dsz[DSZ_NODES];
    for (int c = 0; c < nd.w; c++) {
        const uint32_t cf = (uint32_t)(c < 4 ? rch0[4 * r + c] : rch1[4 * r + c - 4]);
        const uint32_t vc = noff[cf] + (uint32_t)(nd.z + 1 - chain_first_level(K, cf, L));
        if (vc >= nodes) {  // cannot happen for sorted keys; the build is then marked failed
            dsz[DSZ_OVERFLOW] = 1u;
            return;
        }
        child[8 * (size_t)v + c] = (int)vc;
        parent[vc] = (int)v;
    }
}

__device__ inline void kahan(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// FP32 moments, one level per launch (deepest first), 8 lanes per node: the lanes take
// the node's particles (or children) in turn, so a row's loads are consecutive particles, and the
// row's partial sums meet in a butterfly. Each lane sums with Kahan compensation, relative to the
// node's first member (or first child's com); the compensated lane sums are then added in a fixed
// order, so the moments are deterministic. Leaves read their particles; inner nodes read their
// children's moments, which the previous level's launch wrote (and inner nodes of at most
// RMAX_EXACT_MAX particles read their particles again for the exact rmax).
constexpr int MOM_ROW = 8;  // lanes per node: a leaf's <= 32 particles are 4 per lane, an inner node's <= 8 children 1
__device__ inline float row_sum(float v) {
#pragma unroll
    for (int o = MOM_ROW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, MOM_ROW);
    return v;
}
__device__ inline float row_max(float v) {
#pragma unroll
    for (int o = MOM_ROW / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, MOM_ROW));
    return v;
}
// every lane of the node's row calls this with its index `sub` in the row
__device__ inline void moments_f32_node(size_t v, size_t rec, int sub, const float4* __restrict__ pp, const int4* __restrict__ info,
                                        const int* __restrict__ child, float4* __restrict__ com, float4* __restrict__ q0,
                                        float4* __restrict__ q1, float4* __restrict__ wgeo, float4* __restrict__ wq0,
                                        float4* __restrict__ wq1) {
    const int4 nd = info[v];
    float M = 0, cM = 0, Sx = 0, cx = 0, Sy = 0, cy = 0, Sz = 0, cz = 0;
    float Q[6] = {0, 0, 0, 0, 0, 0}, cQ[6] = {0, 0, 0, 0, 0, 0};
    float r2max = 0.f, rbound = 0.f;
    float3 cm;
    if (nd.z == 0) {
        const float4 r = pp[nd.x];
        for (int p = nd.x + sub; p < nd.x + nd.y; p += MOM_ROW) {
            const float4 a = pp[p];
            kahan(M, cM, a.w);
            kahan(Sx, cx, a.w * (a.x - r.x));
            kahan(Sy, cy, a.w * (a.y - r.y));
            kahan(Sz, cz, a.w * (a.z - r.z));
        }
        M = row_sum(M - cM);
        cm = make_float3(r.x + row_sum(Sx - cx) / M, r.y + row_sum(Sy - cy) / M, r.z + row_sum(Sz - cz) / M);
        for (int p = nd.x + sub; p < nd.x + nd.y; p += MOM_ROW) {
            const float4 a = pp[p];
            const float dx = a.x - cm.x, dy = a.y - cm.y, dz = a.z - cm.z;
            kahan(Q[0], cQ[0], a.w * dx * dx);
            kahan(Q[1], cQ[1], a.w * dy * dy);
            kahan(Q[2], cQ[2], a.w * dz * dz);
            kahan(Q[3], cQ[3], a.w * dx * dy);
            kahan(Q[4], cQ[4], a.w * dx * dz);
            kahan(Q[5], cQ[5], a.w * dy * dz);
            r2max = fmaxf(r2max, dx * dx + dy * dy + dz * dz);
        }
    } else {
        // lane c < nd.z takes child c
        const bool has = sub < nd.z;
        const int cv = has ? child[8 * v + sub] : 0;
        const float4 r = com[child[8 * v]];
        const float4 a = has ? com[cv] : make_float4(0.f, 0.f, 0.f, 0.f);
        M = row_sum(a.w);
        cm = make_float3(r.x + row_sum(a.w * (a.x - r.x)) / M, r.y + row_sum(a.w * (a.y - r.y)) / M,
                         r.z + row_sum(a.w * (a.z - r.z)) / M);
        if (has) {
            const float4 b0 = q0[cv], b1 = q1[cv];
            const float dx = a.x - cm.x, dy = a.y - cm.y, dz = a.z - cm.z;
            Q[0] = b0.x + a.w * dx * dx;
            Q[1] = b0.y + a.w * dy * dy;
            Q[2] = b0.z + a.w * dz * dz;
            Q[3] = b0.w + a.w * dx * dy;
            Q[4] = b1.x + a.w * dx * dz;
            Q[5] = b1.y + a.w * dy * dz;
            rbound = sqrtf(dx * dx + dy * dy + dz * dz) + b1.w;
        }
        rbound = row_max(rbound);
        if (nd.y <= RMAX_EXACT_MAX) {
            for (int p = nd.x + sub; p < nd.x + nd.y; p += MOM_ROW) {
                const float4 s = pp[p];
                const float dx = s.x - cm.x, dy = s.y - cm.y, dz = s.z - cm.z;
                r2max = fmaxf(r2max, dx * dx + dy * dy + dz * dz);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) Q[k] = row_sum(Q[k] - cQ[k]);
    r2max = row_max(r2max);
    if (sub == 0) {
        const float rmax = nd.z == 0 || nd.y <= RMAX_EXACT_MAX ? sqrtf(r2max) : rbound;
        com[v] = make_float4(cm.x, cm.y, cm.z, M);
        q0[v] = make_float4(Q[0], Q[1], Q[2], Q[3]);
        q1[v] = make_float4(Q[4], Q[5], Q[0] + Q[1] + Q[2], rmax);
        // the walk's copy at the node's record: what the opening test reads in one load, and the
        // quadrupole with the mass in place of rmax and half the trace (the walk's table holds the
        // trace term's line doubled)
        wgeo[rec] = make_float4(cm.x, cm.y, cm.z, rmax);
        wq0[rec] = make_float4(Q[0], Q[1], Q[2], Q[3]);
        wq1[rec] = make_float4(Q[4], Q[5], 0.5f * (Q[0] + Q[1] + Q[2]), M);
    }
}

constexpr int LEVEL_BLOCKS = 2048;  // grid of the per-level launches (grid-stride over the level)
__global__ void moments_f32_kernel(const float4* __restrict__ pp, const uint32_t* __restrict__ dsz, int level,
                                   const uint32_t* __restrict__ order, const int4* __restrict__ info,
                                   const int* __restrict__ child, float4* __restrict__ com, float4* __restrict__ q0,
                                   float4* __restrict__ q1, float4* __restrict__ wgeo, float4* __restrict__ wq0,
                                   float4* __restrict__ wq1) {
    if (dsz[DSZ_OVERFLOW]) return;
    const uint32_t n = dsz[DSZ_LVL + level], o = dsz[DSZ_LOFF + level];
    const int sub = (int)(threadIdx.x & (MOM_ROW - 1));
    // a node's lanes share i, so they are active or done together and their shuffles stay among them
    for (size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / MOM_ROW; i < n;
         i += (size_t)gridDim.x * blockDim.x / MOM_ROW)
        moments_f32_node(order[o + i], o + i, sub, pp, info, child, com, q0, q1, wgeo, wq0, wq1);
}

// FP64 moments for the plain FP64 path (12 doubles per node).
__device__ inline void moments_f64_node(size_t v, const double* __restrict__ x, const double* __restrict__ y,
                                        const double* __restrict__ z, double mass, const int4* __restrict__ info,
                                        const int* __restrict__ child, double* __restrict__ nd_d) {
    const int4 nd = info[v];
    double M = 0, Sx = 0, Sy = 0, Sz = 0, Q[6] = {0, 0, 0, 0, 0, 0}, rmax = 0;
    double cxx, cyy, czz;
    if (nd.z == 0) {
        for (int p = nd.x; p < nd.x + nd.y; p++) { M += mass; Sx += mass * x[p]; Sy += mass * y[p]; Sz += mass * z[p]; }
        cxx = Sx / M; cyy = Sy / M; czz = Sz / M;
        for (int p = nd.x; p < nd.x + nd.y; p++) {
            double dx = x[p] - cxx, dy = y[p] - cyy, dz = z[p] - czz;
            Q[0] += mass * dx * dx; Q[1] += mass * dy * dy; Q[2] += mass * dz * dz;
            Q[3] += mass * dx * dy; Q[4] += mass * dx * dz; Q[5] += mass * dy * dz;
            rmax = fmax(rmax, sqrt(dx * dx + dy * dy + dz * dz));
        }
    } else {
        for (int c = 0; c < nd.z; c++) {
            const double* a = nd_d + 12 * (size_t)child[8 * v + c];
            M += a[3]; Sx += a[3] * a[0]; Sy += a[3] * a[1]; Sz += a[3] * a[2];
        }
        cxx = Sx / M; cyy = Sy / M; czz = Sz / M;
        for (int c = 0; c < nd.z; c++) {
            const double* a = nd_d + 12 * (size_t)child[8 * v + c];
            double dx = a[0] - cxx, dy = a[1] - cyy, dz = a[2] - czz, m = a[3];
            Q[0] += a[4] + m * dx * dx; Q[1] += a[5] + m * dy * dy; Q[2] += a[6] + m * dz * dz;
            Q[3] += a[7] + m * dx * dy; Q[4] += a[8] + m * dx * dz; Q[5] += a[9] + m * dy * dz;
            rmax = fmax(rmax, sqrt(dx * dx + dy * dy + dz * dz) + a[11]);
        }
        if (nd.y <= RMAX_EXACT_MAX) {
            double r2 = 0;
            for (int p = nd.x; p < nd.x + nd.y; p++) {
                const double dx = x[p] - cxx, dy = y[p] - cyy, dz = z[p] - czz;
                r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
            }
            rmax = sqrt(r2);
        }
    }
    double* o = nd_d + 12 * v;
    o[0] = cxx; o[1] = cyy; o[2] = czz; o[3] = M;
    for (int k = 0; k < 6; k++) o[4 + k] = Q[k];
    o[10] = Q[0] + Q[1] + Q[2];
    o[11] = rmax;
}
__global__ void moments_f64_kernel(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                                   double mass, const uint32_t* __restrict__ dsz, int level, const uint32_t* __restrict__ order,
                                   const int4* __restrict__ info, const int* __restrict__ child, double* __restrict__ nd_d) {
    if (dsz[DSZ_OVERFLOW]) return;
    const uint32_t n = dsz[DSZ_LVL + level], o = dsz[DSZ_LOFF + level];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        moments_f64_node(order[o + i], x, y, z, mass, info, child, nd_d);
}

// Target groups. A node of at most G particles whose parent holds more than G is a group, as is
// a cell root of at most G; a leaf of more than G (coincident particles) splits into chunks of G.
// Consecutive sibling groups are merged while their sum stays <= G: a parent of more than G packs
// its small children, in order, into runs of contiguous particles. The children of a perturbed
// lattice or a halo are uneven (many of a few particles each), so merging cuts the number of
// groups, and with it of tree walks, to about half while keeping each group inside its parent.
// The parent (or the big leaf, or the small cell root) owns its groups; group_pack visits them.
// Run groups (runs == 1): a chaining cell's particles, in their Morton order, are cut into
// consecutive runs of G (the last one shorter), whatever the tree below the cell: every group but
// the last of a cell fills a wavefront.
// Packed groups (runs == 2): tree groups, except that the small children of a parent are packed
// into runs of exactly G consecutive particles (a run may end inside a child), so the runs of a
// parent fill their wavefronts but the last, and each still lies inside its parent.
template <typename F>
__device__ inline uint32_t group_pack(const int4& nd, int L, int G, int runs, const int4* __restrict__ info,
                                      const int* __restrict__ child, const int* __restrict__ parent, size_t v,
                                      F&& emit) {
    const bool packed = runs == 2;
    if (runs == 1) {
        if (nd.w != L) return 0u;
        const uint32_t c = (uint32_t)((nd.y + G - 1) / G);
        for (uint32_t t = 0; t < c; t++) {
            const int b = nd.x + (int)t * G;
            emit(b, min(nd.x + nd.y, b + G) - b);
        }
        return c;
    }
    if (nd.y <= G) {
        // a small node below the cell level belongs to its parent's runs (a node the child table
        // does not link, parent -1, is a group of its own)
        if (nd.w != L && parent[v] >= 0) return 0u;
        emit(nd.x, nd.y);
        return 1u;
    }
    if (nd.z == 0) {
        const uint32_t c = (uint32_t)((nd.y + G - 1) / G);
        for (uint32_t t = 0; t < c; t++) {
            const int b = nd.x + (int)t * G;
            emit(b, min(nd.x + nd.y, b + G) - b);
        }
        return c;
    }
    uint32_t c = 0;
    int rb = 0, rn = 0;  // the open run: first particle, count
    for (int k = 0; k < nd.z; k++) {
        const int4 ci = info[child[8 * v + k]];
        if (ci.y > G) {  // a big child owns its own groups and ends the run
            if (rn) { emit(rb, rn); c++; rn = 0; }
            continue;
        }
        if (packed) {  // the run takes the child's particles and is cut at exactly G
            if (!rn) rb = ci.x;
            rn += ci.y;
            while (rn >= G) { emit(rb, G); c++; rb += G; rn -= G; }
            continue;
        }
        if (rn && rn + ci.y > G) { emit(rb, rn); c++; rn = 0; }
        if (!rn) rb = ci.x;
        rn += ci.y;
    }
    if (rn) { emit(rb, rn); c++; }
    return c;
}
__global__ void group_count_kernel(const uint32_t* __restrict__ dsz, size_t cap, int L, int G, int runs,
                                   const int4* __restrict__ info,
                                   const int* __restrict__ child, const int* __restrict__ parent,
                                   uint32_t* __restrict__ gcnt) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= cap) return;
    if (dsz[DSZ_OVERFLOW] || v >= dev_count(dsz, DSZ_NODES, cap)) {  // entries past the node count scan as zero
        gcnt[v] = 0u;
        return;
    }
    gcnt[v] = group_pack(info[v], L, G, runs, info, child, parent, v, [](int, int) {});
}
__global__ void group_write_kernel(uint32_t* __restrict__ dsz, size_t cap, size_t gcap, int L, int G, int runs,
                                   const int4* __restrict__ info, const int* __restrict__ child,
                                   const int* __restrict__ parent, const uint32_t* __restrict__ gcnt,
                                   const uint32_t* __restrict__ goff,
                                   int2* __restrict__ groups) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dsz[DSZ_OVERFLOW] || v >= dev_count(dsz, DSZ_NODES, cap)) return;
    uint32_t c = gcnt[v];
    if (!c) return;
    if ((size_t)goff[v] + c > gcap) {
        dsz[DSZ_OVERFLOW] = 1u;
        return;
    }
    uint32_t g = goff[v];
    group_pack(info[v], L, G, runs, info, child, parent, v, [&](int b, int m) { groups[g++] = make_int2(b, m); });
}
__global__ void group_box_kernel(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                                 const uint32_t* __restrict__ dsz, size_t gcap, const int2* __restrict__ groups,
                                 double* __restrict__ gbox, int* __restrict__ pgroup) {
    size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dsz[DSZ_OVERFLOW] || g >= dev_count(dsz, DSZ_GROUPS, gcap)) return;
    int2 gr = groups[g];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int p = gr.x; p < gr.x + gr.y; p++) {
        lo[0] = fmin(lo[0], x[p]); hi[0] = fmax(hi[0], x[p]);
        lo[1] = fmin(lo[1], y[p]); hi[1] = fmax(hi[1], y[p]);
        lo[2] = fmin(lo[2], z[p]); hi[2] = fmax(hi[2], z[p]);
        pgroup[p] = (int)g;
    }
    for (int a = 0; a < 3; a++) { gbox[6 * g + a] = lo[a]; gbox[6 * g + 3 + a] = hi[a]; }
}

// start of a tree build: clear the size words and the interaction counters (a kernel, not a
// memset, so the step can be captured in a graph)
__global__ void tree_reset_kernel(uint32_t* __restrict__ dsz, unsigned long long* __restrict__ counters) {
    tree_reset(threadIdx.x, dsz, counters);
}

// ---------------------------------------------------------------------------------------
// Walk helpers
template <int V>
struct ic {  // a compile-time int for generic lambdas
    static constexpr int value = V;
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// periodic image offset of neighbour nb (0..26) of cell (cx,cy,cz)
__device__ inline float image_shift(int c, int d, int mcell, float box) {
    int t = c + d;
    return t < 0 ? -box : (t >= mcell ? box : 0.0f);
}

}  // namespace

// Device view of one box for the walk. The force tables are per unit n^2 (shared by boxes of
// equal law, softening and split), so `fact` carries the box's kick factor times 1/n^2.
struct walk_box {
    const float4* pp;
    // the tree by record (the build's order: the children of a node are consecutive records)
    const int4* link;      // first particle, count, first child record, children
    const float4* geo;     // com, rmax
    const float4* q0;      // Qxx Qyy Qzz Qxy
    const float4* q1;      // Qxz Qyz trQ, mass
    const int* cellroot;   // root record of each chaining cell, or -1
    const int2* groups;
    const uint64_t* key;
    double *tvx, *tvy, *tvz;       // the tree kick of each particle (written, not added: see below)
    unsigned long long* counters;  // [0] n_pp, [1] n_pc, [2] walk stack overflow
    uint32_t* dsz;                 // sizes; the walk sets DSZ_WALKOVF on a stack overflow
    int cell_shift;
    int mcell;
    int naccept;
    float box;       // n
    float rc2;
    float rcut;
    float inv_theta2;
    float fact;
};

namespace {

// Group prefix over the boxes of one walk launch (a box whose build overflowed contributes no
// groups and so gets no kick) and the reset of the group tickets (one per XCD range).
__global__ void walk_prefix_kernel(const walk_box* __restrict__ B, int nbox, uint32_t* __restrict__ prefix,
                                   unsigned long long* __restrict__ ticket) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t s = 0;
    for (int b = 0; b < nbox; b++) {
        prefix[b] = s;
        const uint32_t* d = B[b].dsz;
        s += d[DSZ_OVERFLOW] ? 0u : d[DSZ_GROUPS];
    }
    prefix[nbox] = s;
    for (int x = 0; x < WALK_XCDS; x++) ticket[x] = 0ull;
}

// CENSUS: also count the evaluated particle pairs that lie within the cutoff (counters[3]), a
// measurement of the lists' useful fraction; it adds a compare and an add per pair.
// Pointers read from a descriptor in memory are generic to the compiler, which then emits flat
// loads and stores. A flat access also counts against the LDS wait counter, so every LDS wait that
// follows one waits for the memory access too; these casts make them global ones.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* gld(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T* gst(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
// (the HIP vector types do not copy from a global reference: load as clang vectors)
typedef float gv4f __attribute__((ext_vector_type(4)));
typedef int gv4i __attribute__((ext_vector_type(4)));
typedef int gv2i __attribute__((ext_vector_type(2)));
__device__ inline float4 gld4(const float4* p, size_t i) {
    const gv4f v = ((const __attribute__((address_space(1))) gv4f*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline int4 gld4(const int4* p, size_t i) {
    const gv4i v = ((const __attribute__((address_space(1))) gv4i*)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ inline int2 gld2(const int2* p, size_t i) {
    const gv2i v = ((const __attribute__((address_space(1))) gv2i*)p)[i];
    return make_int2(v.x, v.y);
}

template <bool LOGT, bool CENSUS>
__global__ void __launch_bounds__(WALK_THREADS) walk_f32_kernel(const walk_box* __restrict__ B, const uint32_t* __restrict__ gprefix,
                                                                int nbox, const float2* __restrict__ tab0,
                                                                const float4* __restrict__ tab12, int* __restrict__ stacks,
                                                                unsigned long long* __restrict__ ticket_ctr) {
    // positions are in scaled units (tree_prep_kernel): for the log tables r^2 here is lam r^2
    constexpr float lam = 1.0f;
    // One 1024-thread workgroup per CU: the 16 waves share one LDS copy of the particle force
    // table (32 KB) next to their private interaction lists.
    __shared__ float2 s_tab0[TAB_LDS];
    __shared__ uint32_t s_ppidx[WALK_WAVES][PP_CAP];
    __shared__ int s_stack[WALK_WAVES][LSTACK > 0 ? LSTACK : 1];
    __shared__ int s_pcidx[WALK_WAVES][PC_CAP];
    __shared__ float4 s_pp[WALK_WAVES][64];
    __shared__ float4 s_pc[WALK_WAVES][3][64];
    __shared__ float4 s_shift[WALK_WAVES][27];  // the group's image shift of each neighbour cell
    constexpr int NBIN = LOGT ? LAGPM_LOG_BINS : 4096;
    for (int i = threadIdx.x; i < NBIN; i += WALK_THREADS) s_tab0[i] = tab0[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t* ppidx = s_ppidx[wid];
    int* pcidx = s_pcidx[wid];
    float4* cpp = s_pp[wid];
    float4* cpc0 = s_pc[wid][0];
    float4* cpc1 = s_pc[wid][1];
    float4* cpc2 = s_pc[wid][2];
    int* gstack = stacks + (size_t)(blockIdx.x * WALK_WAVES + wid) * STACK_MAX;
    int* lstack = s_stack[wid];
    // stack entry i: LDS below LSTACK (96% of pop rounds stay there), global memory above
    // (the LDS read is unconditional, from a clamped index, so that the compiler does not merge
    // the two reads into one flat load through a selected pointer: a flat load also holds up every
    // LDS wait behind it)
    auto sget = [&](int i) -> int {
        const int l = lstack[i < LSTACK ? i : LSTACK - 1];
        int g = 0;
        if (i >= LSTACK) g = gstack[i];
        return i < LSTACK ? l : g;
    };
    auto sset = [&](int i, int v) {
        if (i < LSTACK) lstack[i] = v;
        else gstack[i] = v;
    };
    const int gtotal = __builtin_amdgcn_readfirstlane((int)gprefix[nbox]);
    // XCD-local group ranges. Workgroups are dispatched to the XCDs round robin (block b runs on
    // XCD b % 8), and each XCD has its own L2. The groups are in Morton order, so giving XCD x
    // the contiguous range x of 8 keeps the sources and nodes that neighbouring groups share in
    // one L2; a single global ticket would spread neighbours over all eight caches. An XCD that
    // has finished its range takes tickets from the next ranges in turn (work stealing), and a
    // wave leaves after finding all eight ranges exhausted.
    auto range_lo = [&](int x) -> int { return (int)(((long long)gtotal * x) / WALK_XCDS); };
    int slot = (int)(blockIdx.x % WALK_XCDS), tries = 0;

    for (;;) {
        // Group ticket: one atomic per group, from lane 0 (an atomic executed by all 64 lanes is
        // 64 serialised operations on one address, which capped the group rate). The ticket is
        // then read from lane 0 into a scalar register, so every branch of the loop below is
        // wave-uniform: the compiler cannot thread the back edge differently per lane (which
        // would let lanes 1..63 re-enter a group without a new ticket).
        unsigned long long ticket = 0ull;
        if (lane == 0) ticket = atomicAdd(ticket_ctr + slot, 1ull);
        const long long t = (long long)(unsigned)__builtin_amdgcn_readlane((int)ticket, 0) |
                            ((long long)__builtin_amdgcn_readlane((int)(ticket >> 32), 0) << 32);
        const int lo = range_lo(slot), hi = range_lo(slot + 1);
        if (t >= (long long)(hi - lo)) {
            if (++tries == WALK_XCDS) break;
            slot = (slot + 1) % WALK_XCDS;
            continue;
        }
        const int g = lo + (int)t;
        // the box that owns global group g: largest b with gprefix[b] <= g (wave-uniform search)
        int blo = 0, bhi = nbox;
        while (bhi - blo > 1) {
            const int mid = (blo + bhi) >> 1;
            if (__builtin_amdgcn_readfirstlane((int)gprefix[mid]) <= g) blo = mid;
            else bhi = mid;
        }
        const walk_box A = B[blo];
        const int2 gr = gld2(A.groups, g - __builtin_amdgcn_readfirstlane((int)gprefix[blo]));
        const int cnt = __builtin_amdgcn_readfirstlane(gr.y);
        const int gbeg = __builtin_amdgcn_readfirstlane(gr.x);
        // Lane layout. Lane (tgt, rep) holds M targets, tgt + i Tw for i < M, and takes every R-th
        // source of a chunk from rep on; the R partial sums of a target are added at the end.
        // With M = 1, R = 64/Tw lanes per target, Tw the smallest power of two >= cnt, the lanes
        // past cnt idle: for 33-48 targets a third of the wave or more. M = 3 targets per lane
        // fill ceil(cnt/3) slots instead, and each staged source or node is read once for three
        // targets. A group takes M = 3 where it uses at least 4/3 of the lanes M = 1 would
        // (cnt in 33-48, 17-24, 9-12, 5-6), and for 65-192 targets, which one target per lane
        // cannot hold. (Two and five targets per lane as well, for the remaining gaps, made the
        // kernel's code four times larger and spilled registers: the walk got slower.)
        int R = 1;
        while (cnt * 2 * R <= 64) R <<= 1;
        int M = 1;
        {
            const int s3 = (cnt + 2) / 3;
            int R3 = 1;
            while (s3 * 2 * R3 <= 64) R3 <<= 1;
            if (cnt > 64 || R3 >= 4 * R) {
                M = 3;
                R = R3;
            }
        }
        const int Tw = 64 / R;
        const int tgt = lane & (Tw - 1), rep = lane / Tw;
        const uint64_t cellkey = gld(A.key)[gbeg] >> A.cell_shift;
        const int cx = (int)compact3(cellkey), cy = (int)compact3(cellkey >> 1), cz = (int)compact3(cellkey >> 2);
        // the rest of the group, compiled once per targets-per-lane count so that the single-target
        // path carries no registers for the other two
        auto group_body = [&](auto mc) {
        constexpr int M = decltype(mc)::value;
        bool valid[M];
        int me[M];
        float4 xt[M];
        float ax[M], ay[M], az[M];
        float lx = 0.f, ly = 0.f, lz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
#pragma unroll
        for (int i = 0; i < M; i++) {
            const int t = tgt + i * Tw;
            valid[i] = t < cnt;
            me[i] = gbeg + (valid[i] ? t : 0);  // an idle slot repeats target 0: inside the group box
            xt[i] = gld4(A.pp, me[i]);
            ax[i] = ay[i] = az[i] = 0.f;
            lx = i ? fminf(lx, xt[i].x) : xt[i].x; hx = i ? fmaxf(hx, xt[i].x) : xt[i].x;
            ly = i ? fminf(ly, xt[i].y) : xt[i].y; hy = i ? fmaxf(hy, xt[i].y) : xt[i].y;
            lz = i ? fminf(lz, xt[i].z) : xt[i].z; hz = i ? fmaxf(hz, xt[i].z) : xt[i].z;
        }
        const float lox = wave_minf(lx), loy = wave_minf(ly), loz = wave_minf(lz);
        const float hix = wave_maxf(hx), hiy = wave_maxf(hy), hiz = wave_maxf(hz);
        unsigned long long npp = 0, npc = 0;
        unsigned nin = 0;  // CENSUS: this lane's pairs within the cutoff
        int sp = 0;
        bool overflow = false;
        {
            int e = -1;
            if (lane < 27) {
                int dx = lane % 3 - 1, dy = (lane / 3) % 3 - 1, dz = lane / 9 - 1;
                int nx = (cx + dx + A.mcell) % A.mcell, ny = (cy + dy + A.mcell) % A.mcell, nz = (cz + dz + A.mcell) % A.mcell;
                int r = gld(A.cellroot)[nx + A.mcell * (ny + A.mcell * nz)];
                if (r >= 0) e = (r << 5) | lane;
            }
            int tot;
            int off = wave_excl_count(e >= 0, tot);
            if (e >= 0) sset(off, e);
            sp = tot;
        }
        int pp_n = 0, pc_n = 0;

        // The periodic image shift of each of the 27 neighbour cells, made once per group (lane nb
        // makes entry nb) and then read from LDS: a lookup, where computing it (three compares and
        // selects per axis and the base-3 digits of nb) cost some 25 operations per use, and every
        // staged particle, staged node and popped node uses one.
        if (lane < 27)
            s_shift[wid][lane] = make_float4(image_shift(cx, lane % 3 - 1, A.mcell, A.box),
                                             image_shift(cy, (lane / 3) % 3 - 1, A.mcell, A.box),
                                             image_shift(cz, lane / 9 - 1, A.mcell, A.box), 0.f);
        wave_sync();
        auto shift_of = [&](int nb, float& sx, float& sy, float& sz) {
            const float4 sh = s_shift[wid][nb];
            sx = sh.x;
            sy = sh.y;
            sz = sh.z;
        };
        // Source chunks are staged in LDS as float4 (x, y, z, m); entries past m are far-away
        // dummies of zero mass, so every lane runs the same unrolled trip count and the cutoff
        // is a select (no branch, no exec-mask loop).
        // (8 box lengths away: sources lie within one box length of [0, n), and n >= 4 cutoffs)
        const float4 dummy = make_float4(8.f * A.box, 8.f * A.box, 8.f * A.box, 0.f);
        auto load_pp = [&](int start, int m) -> float4 {
            float4 s = dummy;
            if (lane < m) {
                uint32_t e = ppidx[start + lane];
                float sx, sy, sz;
                shift_of((int)(e & 31u), sx, sy, sz);
                s = gld4(A.pp, e >> 5);
                s.x += sx; s.y += sy; s.z += sz;
            }
            return s;
        };
        // Pair sums over one staged chunk. Lane (tgt, rep) takes the sources rep, rep + R, ...
        // (neighbouring lanes read neighbouring entries: no LDS bank conflicts). R is a template
        // value here, so the loads of one round are immediate offsets from one address. Sources
        // are taken four at a time: their LDS loads, then their table loads, are issued together
        // so one wave's LDS latency overlaps its own arithmetic.
        auto pp_one = [&](float4 s) {
#pragma unroll
            for (int i = 0; i < M; i++) {
                float dx = xt[i].x - s.x, dy = xt[i].y - s.y, dz = xt[i].z - s.z;
                float r2 = dx * dx + dy * dy + dz * dz;
                if constexpr (CENSUS) nin += (valid[i] && r2 <= A.rc2) ? 1u : 0u;
                float t;
                const int nt = table_bin<LOGT>(r2, t, lam);
                float2 tb = s_tab0[nt];
                float f = tb.x + tb.y * t;  // particle masses are 1 here: the box mass is in fact
                if constexpr (!LOGT) f = (r2 <= A.rc2) ? f : 0.f;
                ax[i] += dx * f; ay[i] += dy * f; az[i] += dz * f;
            }
        };
        auto pp_chunk_r = [&](auto rc) {
            constexpr int RR = decltype(rc)::value;
            constexpr int iters = 64 / RR;
            constexpr int U = M == 1 ? 4 : 2;  // sources per round: U M pairs in flight
            const float4* src = cpp + rep;
            if constexpr (iters >= U) {
#pragma unroll 2
                for (int k = 0; k < iters; k += U) {
                    float4 s[U];
                    float dx[U][M], dy[U][M], dz[U][M], r2[U][M];
                    bool any = false;
#pragma unroll
                    for (int u = 0; u < U; u++) s[u] = src[(k + u) * RR];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int i = 0; i < M; i++) {
                            dx[u][i] = xt[i].x - s[u].x; dy[u][i] = xt[i].y - s[u].y; dz[u][i] = xt[i].z - s[u].z;
                            r2[u][i] = dx[u][i] * dx[u][i] + dy[u][i] * dy[u][i] + dz[u][i] * dz[u][i];
                            if constexpr (!LOGT) any = any || (r2[u][i] <= A.rc2);
                            if constexpr (CENSUS) nin += (valid[i] && r2[u][i] <= A.rc2) ? 1u : 0u;
                        }
                    if constexpr (!LOGT)
                        if (!__any(any)) continue;  // wave-uniform skip of U far sources
                    float2 tb[U][M];
                    float t[U][M];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int i = 0; i < M; i++) tb[u][i] = s_tab0[table_bin<LOGT>(r2[u][i], t[u][i], lam)];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int i = 0; i < M; i++) {
                            float f = tb[u][i].x + tb[u][i].y * t[u][i];
                            if constexpr (!LOGT) f = (r2[u][i] <= A.rc2) ? f : 0.f;
                            ax[i] += dx[u][i] * f; ay[i] += dy[u][i] * f; az[i] += dz[u][i] * f;
                        }
                }
            } else {
#pragma unroll
                for (int k = 0; k < iters; k++) pp_one(src[k * RR]);
            }
        };
        // the common replications take the unrolled loops (one target per lane: groups of more
        // than 8 targets, R <= 4; three: more than 12, R <= 8); the rest a plain loop
        auto pp_chunk = [&]() {
            if (R == 1) pp_chunk_r(ic<1>{});
            else if (R == 2) pp_chunk_r(ic<2>{});
            else if (R == 4) pp_chunk_r(ic<4>{});
            else if (M == 3 && R == 8) pp_chunk_r(ic<8>{});
            else
                for (int k = 0; k < 64 / R; k++) pp_one(cpp[rep + k * R]);
        };
        // flush `nchunk` consecutive chunks starting at list entry `start` (the last may be
        // partial: `mlast` entries); the next chunk's global loads overlap this chunk's math
        // Staging with a cut: a listed particle farther than the cutoff from the group's bounding
        // box is beyond it from every target, so its pair terms are exactly zero (the table is zero
        // there, and the reference law's select zeroes it); it is dropped before it costs a chunk
        // slot. The margin keeps every dropped pair clear of the float-rounded table edge.
        const float rc2_keep = A.rc2 * (1.0f + 1.0e-5f);
        auto near_group = [&](float4 q) -> bool {
            const float ddx = fmaxf(fmaxf(lox - q.x, q.x - hix), 0.f);
            const float ddy = fmaxf(fmaxf(loy - q.y, q.y - hiy), 0.f);
            const float ddz = fmaxf(fmaxf(loz - q.z, q.z - hiz), 0.f);
            return ddx * ddx + ddy * ddy + ddz * ddz <= rc2_keep;
        };
        int fill = 0;  // staged particles in cpp[0, fill), wave-uniform
        // append the kept particles of one 64-entry block to the chunk, computing the chunk when it
        // fills; survivors past its end start the next chunk
        auto stage_pp = [&](float4 q, bool keep) {
            int tot;
            const int off = wave_excl_count(keep, tot);
            const int room = 64 - fill;
            if (keep && off < room) cpp[fill + off] = q;
            if (tot >= room) {
                wave_sync();
                pp_chunk();
                npp += 64ull;
                wave_sync();
                if (keep && off >= room) cpp[off - room] = q;
                fill = tot - room;
            } else {
                fill += tot;
            }
            wave_sync();
        };
        // stage `nblock` consecutive blocks of the list from entry `start` (the last may be
        // partial: `mlast` entries); the next block's global loads overlap this block's chunk
        auto flush_pp = [&](int start, int nblock, int mlast) {
            float4 nxt = load_pp(start, nblock == 1 ? mlast : 64);
            for (int c = 0; c < nblock; c++) {
                const float4 cur = nxt;
                if (c + 1 < nblock) nxt = load_pp(start + 64 * (c + 1), c + 2 == nblock ? mlast : 64);
                stage_pp(cur, near_group(cur));  // entries past m are far dummies: never kept
            }
        };
        // One node interaction: monopole plus quadrupole, Q d computed once and reused for
        // d.Q.d (the quadratic form) and for the 2 f2 Q d term. The table's f2 line is stored
        // doubled and the node's trace halved (both exact), so 2 f2 is the table value itself.
        auto pc_eval = [&](float4 c, float4 qa, float4 qb, float2 t0, float4 t12, float t, float dx, float dy, float dz,
                           bool in, float& ax, float& ay, float& az) {
            float f1 = t0.x + t0.y * t, f2 = t12.x + t12.y * t, f3 = t12.z + t12.w * t;
            float qx1 = qa.x * dx + qa.w * dy + qb.x * dz;
            float qx2 = qa.w * dx + qa.y * dy + qb.y * dz;
            float qx3 = qb.x * dx + qb.y * dy + qa.z * dz;
            float qxy = dx * qx1 + dy * qx2 + dz * qx3;
            float tmp = c.w * f1 + qb.z * f2 + qxy * f3;
            float two = f2;
            if constexpr (!LOGT) {
                tmp = in ? tmp : 0.f;
                two = in ? two : 0.f;
            }
            ax = fmaf(qx1, two, fmaf(dx, tmp, ax));
            ay = fmaf(qx2, two, fmaf(dy, tmp, ay));
            az = fmaf(qx3, two, fmaf(dz, tmp, az));
        };
        // one node against this lane's M targets
        auto pc_node = [&](const float4& c, const float4& qa, const float4& qb) {
#pragma unroll
            for (int i = 0; i < M; i++) {
                const float dx = xt[i].x - c.x, dy = xt[i].y - c.y, dz = xt[i].z - c.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                float t;
                const int nt = table_bin<LOGT>(r2, t, lam);
                pc_eval(c, qa, qb, s_tab0[nt], tab12[nt], t, dx, dy, dz, r2 <= A.rc2, ax[i], ay[i], az[i]);
            }
        };
        auto pc_chunk_r = [&](auto rc) {
            constexpr int RR = decltype(rc)::value;
            constexpr int iters = 64 / RR;
            constexpr int U = M == 1 ? 2 : 1;  // nodes per round: U M interactions in flight
            const float4* n0 = cpc0 + rep;
            const float4* n1 = cpc1 + rep;
            const float4* n2 = cpc2 + rep;
            if constexpr (iters >= 2) {
                // the nodes' loads first, then their table loads, as for the particles
#pragma unroll 2
                for (int k = 0; k < iters; k += U) {
                    float4 c[U], qa[U], qb[U];
                    float dx[U][M], dy[U][M], dz[U][M], r2[U][M];
                    bool any = false;
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int j = (k + u) * RR;
                        c[u] = n0[j]; qa[u] = n1[j]; qb[u] = n2[j];
                    }
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int i = 0; i < M; i++) {
                            dx[u][i] = xt[i].x - c[u].x; dy[u][i] = xt[i].y - c[u].y; dz[u][i] = xt[i].z - c[u].z;
                            r2[u][i] = dx[u][i] * dx[u][i] + dy[u][i] * dy[u][i] + dz[u][i] * dz[u][i];
                            if constexpr (!LOGT) any = any || (r2[u][i] <= A.rc2);
                        }
                    if constexpr (!LOGT)
                        if (!__any(any)) continue;  // wave-uniform skip of far nodes
                    float t[U][M];
                    float2 t0[U][M];
                    float4 t12[U][M];
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int i = 0; i < M; i++) {
                            const int nt = table_bin<LOGT>(r2[u][i], t[u][i], lam);
                            t0[u][i] = s_tab0[nt];
                            t12[u][i] = tab12[nt];
                        }
#pragma unroll
                    for (int u = 0; u < U; u++)
#pragma unroll
                        for (int i = 0; i < M; i++)
                            pc_eval(c[u], qa[u], qb[u], t0[u][i], t12[u][i], t[u][i], dx[u][i], dy[u][i], dz[u][i],
                                    r2[u][i] <= A.rc2, ax[i], ay[i], az[i]);
                }
            } else {
                pc_node(n0[0], n1[0], n2[0]);
            }
        };
        auto flush_pc = [&](int start, int m) {
            if (lane < m) {
                int e = pcidx[start + lane];
                float sx, sy, sz;
                shift_of(e & 31, sx, sy, sz);
                int v = e >> 5;
                const float4 c = gld4(A.geo, v), q1 = gld4(A.q1, v);
                cpc0[lane] = make_float4(c.x + sx, c.y + sy, c.z + sz, q1.w);  // (centre, mass)
                cpc1[lane] = gld4(A.q0, v);
                cpc2[lane] = q1;
            } else {
                cpc0[lane] = dummy;
                cpc1[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
                cpc2[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            wave_sync();
            if (R == 1) pc_chunk_r(ic<1>{});
            else if (R == 2) pc_chunk_r(ic<2>{});
            else if (R == 4) pc_chunk_r(ic<4>{});
            else if (M == 3 && R == 8) pc_chunk_r(ic<8>{});
            else
                for (int k = 0; k < 64 / R; k++) {
                    const int j = rep + k * R;
                    pc_node(cpc0[j], cpc1[j], cpc2[j]);
                }
            npc += (unsigned long long)m;
            wave_sync();
        };

        int iters = 0;
        while (sp > 0) {
            if (++iters > WALK_MAX_ITERS) {  // a well-formed tree never gets here
                overflow = true;
                break;
            }
            const int take = min(sp, 64);
            sp -= take;
            int e = lane < take ? sget(sp + lane) : -1;
            int kind = 0;  // 0 none, 1 accept, 2 leaf, 3 open
            int4 nd = make_int4(0, 0, 0, 0);  // (first particle, count, first child record, children)
            int v = 0, nb = 0;
            if (e >= 0) {
                v = e >> 5;
                nb = e & 31;
                float sx, sy, sz;
                shift_of(nb, sx, sy, sz);
                // two 16-byte loads per node: the children are consecutive records, so opening a
                // node needs only the first one's index, and the popped nodes of a round (mostly
                // the children of the previous round's opened nodes) share cache lines
                const float4 c = gld4(A.geo, v);
                const float rmax = c.w;
                nd = gld4(A.link, v);
                float px = c.x + sx, py = c.y + sy, pz = c.z + sz;
                float ddx = fmaxf(fmaxf(lox - px, px - hix), 0.f);
                float ddy = fmaxf(fmaxf(loy - py, py - hiy), 0.f);
                float ddz = fmaxf(fmaxf(loz - pz, pz - hiz), 0.f);
                float dmin2 = ddx * ddx + ddy * ddy + ddz * ddz;
                float rr = A.rcut + rmax;
                if (dmin2 > rr * rr) kind = 0;
                else if (nd.y >= A.naccept && dmin2 >= rmax * rmax * A.inv_theta2) kind = 1;
                else if (nd.w == 0) kind = 2;
                else kind = 3;
            }
            // children onto the stack
            {
                int nch = kind == 3 ? nd.w : 0;
                int tot;
                int off = wave_excl_scan(nch, lane, tot);
                if (sp + tot > STACK_MAX) {
                    overflow = true;
                    break;
                }
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (c < nch) sset(sp + off + c, ((nd.z + c) << 5) | nb);
                sp += tot;
                wave_sync();
            }
            // accepted nodes
            {
                int tot;
                int off = wave_excl_count(kind == 1, tot);
                if (kind == 1) pcidx[pc_n + off] = e;
                pc_n += tot;
                wave_sync();
                if (pc_n >= 64) {
                    flush_pc(0, 64);
                    pc_n -= 64;
                    int t = (lane < pc_n) ? pcidx[64 + lane] : 0;
                    wave_sync();
                    if (lane < pc_n) pcidx[lane] = t;
                    wave_sync();
                }
            }
            // opened leaves: their particles
            {
                int rem = kind == 2 ? nd.y : 0;
                int pb = nd.x;
                for (;;) {
                    int tk = min(rem, PP_TAKE);
                    int tot;
                    int off = wave_excl_scan(tk, lane, tot);
                    if (tot == 0) break;
                    if (pp_n + tot > PP_CAP) {
                        int full = pp_n & ~63;
                        if (full) flush_pp(0, full / 64, 64);
                        int left = pp_n - full;
                        uint32_t t = (lane < left) ? ppidx[full + lane] : 0u;
                        wave_sync();
                        if (lane < left) ppidx[lane] = t;
                        wave_sync();
                        pp_n = left;
                    }
                    for (int t = 0; t < tk; t++) ppidx[pp_n + off + t] = ((uint32_t)(pb + t) << 5) | (uint32_t)nb;
                    pp_n += tot;
                    pb += tk;
                    rem -= tk;
                    wave_sync();
                }
                if (pp_n >= 64) {
                    int full = pp_n & ~63;
                    flush_pp(0, full / 64, 64);
                    int left = pp_n - full;
                    uint32_t t = (lane < left) ? ppidx[full + lane] : 0u;
                    wave_sync();
                    if (lane < left) ppidx[lane] = t;
                    wave_sync();
                    pp_n = left;
                }
            }
        }
        if (pc_n > 0) flush_pc(0, pc_n);
        if (pp_n > 0) flush_pp(0, 1, pp_n);
        if (fill > 0) {  // the last, partial chunk: far zero-mass dummies after the staged particles
            if (lane >= fill) cpp[lane] = dummy;
            wave_sync();
            pp_chunk();
            npp += (unsigned long long)fill;
            wave_sync();
        }
        // sum the R partial accumulators of each target (lanes tgt, tgt + Tw, ...)
        for (int o = Tw; o < 64; o <<= 1) {
#pragma unroll
            for (int i = 0; i < M; i++) {
                ax[i] += __shfl_xor(ax[i], o, 64);
                ay[i] += __shfl_xor(ay[i], o, 64);
                az[i] += __shfl_xor(az[i], o, 64);
            }
        }
        // The kick is written to its own arrays rather than added to the velocities, so the walk
        // can run while the PM kick of the same step updates them; the drift (or, for a tree
        // kick alone, an add kernel) adds it afterwards, in the same order as v += pm; v += tree.
        if (overflow && lane == 0) {
            atomicAdd(&A.counters[2], 1ull);
            atomicOr(&A.dsz[DSZ_WALKOVF], 1u);  // the step's drift reads it and leaves the box unmoved
        }
#pragma unroll
        for (int i = 0; i < M; i++) {
            if (valid[i] && rep == 0) {
                gst(A.tvx)[me[i]] = overflow ? 0.0 : (double)(A.fact * ax[i]);
                gst(A.tvy)[me[i]] = overflow ? 0.0 : (double)(A.fact * ay[i]);
                gst(A.tvz)[me[i]] = overflow ? 0.0 : (double)(A.fact * az[i]);
            }
        }
        if (lane == 0) {
            // (global atomics: a flat one would hold up the next group's first LDS wait)
            __hip_atomic_fetch_add(gst(A.counters), npp * (unsigned long long)cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gst(A.counters) + 1, npc * (unsigned long long)cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if constexpr (CENSUS) {
            for (int o = 32; o > 0; o >>= 1) nin += (unsigned)__shfl_xor((int)nin, o, 64);
            if (lane == 0) __hip_atomic_fetch_add(gst(A.counters) + 3, (unsigned long long)nin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        };
        if (M == 3) group_body(ic<3>{});
        else group_body(ic<1>{});
    }
}

// Plain FP64 walk: one thread per target, private stack, same tree, groups and criteria.
struct walk64_args {
    const double *x, *y, *z;
    double mass;
    const int4* info;
    const int* child;
    const double* nd;
    const int* cellroot;
    const uint64_t* key;
    int cell_shift;
    const double* gbox;
    const int* pgroup;
    const double* tab;
    float lam;
    double *vx, *vy, *vz;
    unsigned long long* counters;
    uint32_t* dsz;
    size_t np;
    int mcell;
    double box, cellw, rc2, rcut, inv_theta2, fact;
    int naccept;
    // sampled walk: target t is the particle at sample[t].x, opened as a group of its own (the
    // per-particle test), its acceleration written to out[3t..3t+2] (velocities untouched)
    const int2* sample;
    double* out;
    size_t nsample;
};

template <bool LOGT>
// one target particle; returns false on a stack overflow (the kick is then not applied)
__device__ inline bool walk_f64_target(const walk64_args& A, size_t i, unsigned long long& npp, unsigned long long& npc,
                                       double* out) {
    const double px0 = A.x[i], py0 = A.y[i], pz0 = A.z[i];
    double lo[3] = {px0, py0, pz0}, hi[3] = {px0, py0, pz0};  // a group of one (sampled walk)
    if (!out) {
        const int g = A.pgroup[i];
        for (int a = 0; a < 3; a++) {
            lo[a] = A.gbox[6 * g + a];
            hi[a] = A.gbox[6 * g + 3 + a];
        }
    }
    // all group members share a chaining cell: the level-L prefix of their Morton key
    const uint64_t cellkey = A.key[i] >> A.cell_shift;
    const int cx = (int)compact3(cellkey), cy = (int)compact3(cellkey >> 1), cz = (int)compact3(cellkey >> 2);
    int stack[256];
    int sp = 0;
    for (int nb = 26; nb >= 0; nb--) {
        int dx = nb % 3 - 1, dy = (nb / 3) % 3 - 1, dz = nb / 9 - 1;
        int nx = (cx + dx + A.mcell) % A.mcell, ny = (cy + dy + A.mcell) % A.mcell, nz = (cz + dz + A.mcell) % A.mcell;
        int r = A.cellroot[nx + A.mcell * (ny + A.mcell * nz)];
        if (r >= 0) stack[sp++] = (r << 5) | nb;
    }
    double ax = 0, ay = 0, az = 0;
    bool overflow = false;
    while (sp > 0) {
        int e = stack[--sp];
        int v = e >> 5, nb = e & 31;
        int tx = cx + nb % 3 - 1, ty = cy + (nb / 3) % 3 - 1, tz = cz + nb / 9 - 1;
        double sx = tx < 0 ? -A.box : (tx >= A.mcell ? A.box : 0.0);
        double sy = ty < 0 ? -A.box : (ty >= A.mcell ? A.box : 0.0);
        double sz = tz < 0 ? -A.box : (tz >= A.mcell ? A.box : 0.0);
        const double* m = A.nd + 12 * (size_t)v;
        int4 nd = A.info[v];
        double qx = m[0] + sx, qy = m[1] + sy, qz = m[2] + sz, rmax = m[11];
        double ddx = fmax(fmax(lo[0] - qx, qx - hi[0]), 0.0);
        double ddy = fmax(fmax(lo[1] - qy, qy - hi[1]), 0.0);
        double ddz = fmax(fmax(lo[2] - qz, qz - hi[2]), 0.0);
        double dmin2 = ddx * ddx + ddy * ddy + ddz * ddz;
        double rr = A.rcut + rmax;
        if (dmin2 > rr * rr) continue;
        if (nd.y >= A.naccept && dmin2 >= rmax * rmax * A.inv_theta2) {
            npc++;
            double dx = px0 - qx, dy = py0 - qy, dz = pz0 - qz;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= A.rc2) {
                double t;
                const int nt = table_bin64<LOGT>(r2, t, A.lam);
                const double* tb = A.tab + 6 * nt;
                double f1 = tb[0] + tb[1] * t, f2 = tb[2] + tb[3] * t, f3 = tb[4] + tb[5] * t;
                double qx1 = m[4] * dx + m[7] * dy + m[8] * dz;
                double qx2 = m[7] * dx + m[5] * dy + m[9] * dz;
                double qx3 = m[8] * dx + m[9] * dy + m[6] * dz;
                double qxy = m[4] * dx * dx + m[5] * dy * dy + m[6] * dz * dz + 2 * (m[7] * dx * dy + m[8] * dx * dz + m[9] * dy * dz);
                double tmp = m[3] * f1 + m[10] * f2 + qxy * f3;
                ax += dx * tmp + qx1 * 2 * f2;
                ay += dy * tmp + qx2 * 2 * f2;
                az += dz * tmp + qx3 * 2 * f2;
            }
        } else if (nd.z == 0) {
            for (int p = nd.x; p < nd.x + nd.y; p++) {
                npp++;
                double dx = px0 - (A.x[p] + sx), dy = py0 - (A.y[p] + sy), dz = pz0 - (A.z[p] + sz);
                double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= A.rc2) {
                    double t;
                    const int nt = table_bin64<LOGT>(r2, t, A.lam);
                    double f = A.tab[6 * nt] + A.tab[6 * nt + 1] * t;  // unit particle mass
                    ax += dx * f; ay += dy * f; az += dz * f;
                }
            }
        } else {
            if (sp + nd.z > 256) {
                overflow = true;
                break;
            }
            for (int c = nd.z - 1; c >= 0; c--) stack[sp++] = (A.child[8 * v + c] << 5) | nb;
        }
    }
    if (overflow) return false;
    if (out) {
        out[0] = A.fact * ax;
        out[1] = A.fact * ay;
        out[2] = A.fact * az;
        return true;
    }
    A.vx[i] += A.fact * ax;
    A.vy[i] += A.fact * ay;
    A.vz[i] += A.fact * az;
    return true;
}
__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, o, 64);
        const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), o, 64);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}
// Thread per target. The interaction counters are summed over the wave first: one atomic per
// wave and counter (an atomic per thread on one address serialises millions of operations).
template <bool LOGT>
__global__ void __launch_bounds__(128) walk_f64_kernel(walk64_args A) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long npp = 0, npc = 0, ovf = 0;
    if (!A.dsz[DSZ_OVERFLOW]) {
        if (A.sample) {
            if (t < A.nsample)
                ovf = walk_f64_target<LOGT>(A, (size_t)A.sample[t].x, npp, npc, A.out + 3 * t) ? 0ull : 1ull;
        } else if (t < A.np) {
            ovf = walk_f64_target<LOGT>(A, t, npp, npc, nullptr) ? 0ull : 1ull;
        }
    }
    npp = wave_sum_u64(npp);
    npc = wave_sum_u64(npc);
    ovf = wave_sum_u64(ovf);
    if ((threadIdx.x & 63) == 0) {
        if (npp) atomicAdd(&A.counters[0], npp);
        if (npc) atomicAdd(&A.counters[1], npc);
        if (ovf) {
            atomicAdd(&A.counters[2], ovf);
            atomicOr(&A.dsz[DSZ_WALKOVF], 1u);
        }
    }
}

unsigned blocks_for(size_t n, int t) { return (unsigned)std::max<size_t>(1, (n + t - 1) / t); }

template <typename T>
int regrow(lagpm_ctx* ctx, T** p, size_t count) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    HIPCHK(ctx, hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
    return LAGPM_OK;
}

}  // namespace

int lagpm_choose_cell_level(int n, double rsphere) {
    int L = 0;
    while (L < LAGPM_KEY_BITS && (double)n / (double)(1 << (L + 1)) >= rsphere) L++;
    return L;
}

size_t lagpm_tree_node_cap_guess(size_t np, int mcell) {
    // nleaf = 16 gives about np/6 nodes for uniform and np/4 for strongly clustered input
    return np / 3 + (size_t)mcell * mcell * mcell + 4096;
}

// scale of the FP32 tree's coordinates: sqrt(lam) for the log-r^2 tables, 1 for the r-table law
double lagpm_tree_scale(const lagpm_box* box) {
    return box->tab && box->tab->logr2 ? std::sqrt((double)box->tab->lam) : 1.0;
}

int lagpm_tree_alloc(lagpm_ctx* ctx, lagpm_box* box, size_t cap) {
    const bool f64 = box->p.precision == LAGPM_PREC_FP64;
    if (cap > (size_t)LAGPM_MAX_NODES) {
        if (box->node_cap >= (size_t)LAGPM_MAX_NODES)
            return lagpm_fail(ctx, LAGPM_E_INVAL, "tree needs more than 2^26 nodes (the walk's node index limit)");
        cap = LAGPM_MAX_NODES;
    }
    lagpm_box_graphs_release(box);  // captured steps refer to the old arrays
    const size_t gcap = cap + box->np / 64 + 1;
    RCHK(regrow(ctx, &box->nd_info, cap));
    RCHK(regrow(ctx, &box->nd_parent, cap));
    RCHK(regrow(ctx, &box->nd_child, 8 * cap));
    RCHK(regrow(ctx, &box->gflag, cap));
    RCHK(regrow(ctx, &box->goff, cap));
    RCHK(regrow(ctx, &box->groups, gcap));
    if (f64) {
        RCHK(regrow(ctx, &box->nd_d, 12 * cap));
        RCHK(regrow(ctx, &box->gbox, 6 * gcap));
    } else {
        RCHK(regrow(ctx, &box->nd_com, cap));
        RCHK(regrow(ctx, &box->nd_q0, cap));
        RCHK(regrow(ctx, &box->nd_q1, cap));
        RCHK(regrow(ctx, &box->w_link, cap));
        RCHK(regrow(ctx, &box->w_geo, cap));
        RCHK(regrow(ctx, &box->w_q0, cap));
        RCHK(regrow(ctx, &box->w_q1, cap));
        RCHK(regrow(ctx, &box->w_rfc, cap));
    }
    const size_t need = std::max(lagpm_scan_tmp_bytes(cap), std::max(lagpm_scan_tmp_bytes(box->np), lagpm_sort_tmp_bytes(box->np)));
    if (need > box->scan_tmp_bytes) {
        RCHK(regrow(ctx, (char**)&box->scan_tmp, need));
        box->scan_tmp_bytes = need;
    }
    box->node_cap = cap;
    box->group_cap = gcap;
    return LAGPM_OK;
}

int lagpm_sort_particles(lagpm_ctx* ctx, lagpm_box* box, const double* drag, hipStream_t vst) {
    const size_t np = box->np;
    if (np == 0) return LAGPM_OK;
    hipStream_t st = box->st;
    const double s = 65536.0 / (double)box->p.n;
    hipLaunchKernelGGL(keys_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, box->x, box->y, box->z, np, s,
                       box->key, box->perm);
    HIPCHK(ctx, hipGetLastError());
    // The tree splits nodes down to box->depth at most (deeper trees are detected and redone after
    // a full sort), so only the top 3 * depth key bits must be in order: a box whose tree is
    // shallower needs fewer passes. Below them the order is the previous one (the sort is stable).
    // An even number of passes leaves the result in key/perm, where captured steps and the walk's
    // descriptors expect it.
    const int D = std::min(LAGPM_KEY_BITS, std::max(box->level_cell, box->depth));
    const int passes = (((3 * D + 7) / 8) + 1) & ~1;
    bool alt = false;
    RCHK(lagpm_radix_sort_pairs(ctx, st, box->scan_tmp, box->key, box->key2, box->perm, box->perm2, np, 8 * passes, &alt,
                                3 * LAGPM_KEY_BITS - 8 * passes));
    if (alt) return lagpm_fail(ctx, LAGPM_E_STATE, "radix sort ended in the scratch buffers");
    double** arrs[6] = {&box->x, &box->y, &box->z, &box->vx, &box->vy, &box->vz};
    gather_set g;
    for (int a = 0; a < 6; a++) {
        g.src[a] = *arrs[a];
        g.dst[a] = box->alt[a];
    }
    g.id = box->id;
    g.id_dst = box->id_alt;
    if (vst && vst != st) {
        // the positions first, alone (everything after the sort waits for them), then the
        // velocities and ids on vst beside what follows; the caller joins box->evvel before using them
        hipLaunchKernelGGL(gather_pos_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, g, box->perm, np);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(box->evsort, st));
        HIPCHK(ctx, hipStreamWaitEvent(vst, box->evsort, 0));
        hipLaunchKernelGGL(gather_vel_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, vst, g, box->perm, np, drag);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(box->evvel, vst));
    } else {
        hipLaunchKernelGGL(gather_all_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, g, box->perm, np, drag);
    }
    HIPCHK(ctx, hipGetLastError());
    for (int a = 0; a < 6; a++) std::swap(*arrs[a], box->alt[a]);
    std::swap(box->id, box->id_alt);
    box->parity ^= 1;
    box->sorted = true;
    return LAGPM_OK;
}

int lagpm_tree_build(lagpm_ctx* ctx, lagpm_box* box) {
    const size_t np = box->np;
    const int L = box->level_cell;
    const int nleaf = box->tune.nleaf;
    const int G = box->tune.group;
    const bool f64 = box->p.precision == LAGPM_PREC_FP64;
    hipStream_t st = box->st;
    const size_t cap = box->node_cap, gcap = box->group_cap;
    if (np == 0) {
        hipLaunchKernelGGL(tree_reset_kernel, dim3(1), dim3(64), 0, st, box->dsz, box->counters);
        HIPCHK(ctx, hipGetLastError());
        return LAGPM_OK;
    }
    if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
    // sizes, counters, node counts and cell roots cleared; for the FP32 tree the pp copy in units
    // of the particle mass (the walk multiplies by it with the kick)
    const size_t ncell = (size_t)box->mcell * box->mcell * box->mcell;
    hipLaunchKernelGGL(tree_prep_kernel, dim3(blocks_for(std::max(np, ncell), 256)), dim3(256), 0, st, box->x, box->y,
                       box->z, np, f64 ? 1.0 : lagpm_tree_scale(box), f64 ? (float4*)nullptr : box->pp, box->ncount,
                       box->cellroot, f64 ? (int*)nullptr : box->w_cellroot, ncell, box->dsz, box->counters);
    HIPCHK(ctx, hipGetLastError());
    // top-down build: records of level-L cells, then one split per level, then the numbering.
    // The records and child lists live in the moment arrays until the moments overwrite them.
    int4* rec;
    int *rch0, *rch1;
    if (f64) {
        rec = reinterpret_cast<int4*>(box->nd_d);
        rch0 = reinterpret_cast<int*>(box->nd_d) + 4 * cap;
        rch1 = reinterpret_cast<int*>(box->nd_d) + 8 * cap;
    } else {
        rec = reinterpret_cast<int4*>(box->nd_com);
        rch0 = reinterpret_cast<int*>(box->nd_q0);
        rch1 = reinterpret_cast<int*>(box->nd_q1);
    }
    uint32_t* order = box->goff;  // level order (goff is free until the group scan)
    hipLaunchKernelGGL(cells_kernel, dim3(blocks_for(ncell, 256)), dim3(256), 0, st, box->key, np, L, cap, box->dsz, rec,
                       box->ncount);
    hipLaunchKernelGGL(level_mark_kernel, dim3(1), dim3(64), 0, st, box->dsz, cap, L, L);
    HIPCHK(ctx, hipGetLastError());
    const unsigned lblocks = (unsigned)std::min<size_t>(LEVEL_BLOCKS, blocks_for(cap, 256));
    const unsigned mblocks = (unsigned)std::min<size_t>(2 * LEVEL_BLOCKS, blocks_for(cap * MOM_ROW, 256));
    const int D = std::min(LAGPM_KEY_BITS, std::max(L, box->depth));
    int* rfc = f64 ? (int*)nullptr : box->w_rfc;
    for (int l = L; l < D; l++) {
        hipLaunchKernelGGL(split_kernel, dim3(lblocks), dim3(SPLIT_THREADS), 0, st, box->key, l, D, nleaf, cap, box->dsz,
                           rec, rch0, rch1, box->ncount, rfc);
        hipLaunchKernelGGL(level_mark_kernel, dim3(1), dim3(64), 0, st, box->dsz, cap, l + 1, L);
        HIPCHK(ctx, hipGetLastError());
    }
    if (D < LAGPM_KEY_BITS) {  // the limit level: flag a record that still needs splitting
        hipLaunchKernelGGL(split_kernel, dim3(lblocks), dim3(SPLIT_THREADS), 0, st, box->key, D, D, nleaf, cap, box->dsz,
                           rec, rch0, rch1, box->ncount, rfc);
        HIPCHK(ctx, hipGetLastError());
    }
    RCHK(lagpm_exclusive_scan_u32(ctx, st, box->scan_tmp, box->ncount, box->noff, np, box->dsz + DSZ_SCRATCH));
    hipLaunchKernelGGL(number_kernel, dim3(blocks_for(cap, 256)), dim3(256), 0, st, box->key, L, cap, box->dsz, rec, rch0,
                       rch1, box->noff, box->nd_info, box->nd_child, box->nd_parent, box->cellroot, box->mcell, order, rfc,
                       f64 ? (int4*)nullptr : box->w_link, f64 ? (int*)nullptr : box->w_cellroot);
    HIPCHK(ctx, hipGetLastError());
    for (int l = D; l >= L; l--) {  // levels below D hold no records
        if (f64)
            hipLaunchKernelGGL(moments_f64_kernel, dim3(lblocks), dim3(256), 0, st, box->x, box->y, box->z, 1.0, box->dsz, l,
                               order, box->nd_info, box->nd_child, box->nd_d);
        else  // (MOM_ROW lanes per node: a larger grid keeps as many nodes in flight)
            hipLaunchKernelGGL(moments_f32_kernel, dim3(mblocks), dim3(256), 0, st, box->pp, box->dsz, l, order,
                               box->nd_info, box->nd_child, box->nd_com, box->nd_q0, box->nd_q1, box->w_geo, box->w_q0,
                               box->w_q1);
        HIPCHK(ctx, hipGetLastError());
    }

    const int runs = box->tune.runs;
    hipLaunchKernelGGL(group_count_kernel, dim3(blocks_for(cap, 256)), dim3(256), 0, st, box->dsz, cap, L, G, runs,
                       box->nd_info, box->nd_child, box->nd_parent, box->gflag);
    HIPCHK(ctx, hipGetLastError());
    RCHK(lagpm_exclusive_scan_u32(ctx, st, box->scan_tmp, box->gflag, box->goff, cap, box->dsz + DSZ_GROUPS));
    hipLaunchKernelGGL(group_write_kernel, dim3(blocks_for(cap, 256)), dim3(256), 0, st, box->dsz, cap, gcap, L, G, runs,
                       box->nd_info, box->nd_child, box->nd_parent, box->gflag, box->goff, box->groups);
    HIPCHK(ctx, hipGetLastError());
    if (f64) {
        hipLaunchKernelGGL(group_box_kernel, dim3(blocks_for(gcap, 256)), dim3(256), 0, st, box->x, box->y, box->z,
                           box->dsz, gcap, box->groups, box->gbox, box->pgroup);
        HIPCHK(ctx, hipGetLastError());
    }
    box->tree_built = true;
    return LAGPM_OK;
}

int lagpm_tree_walk_f32(lagpm_ctx* ctx, lagpm_box* const* boxes, const double* facts, int nbox, int reserve_cu,
                        const walk_sample* sample) {
    if (nbox <= 0) return LAGPM_OK;
    if (sample && nbox != 1) return lagpm_fail(ctx, LAGPM_E_INVAL, "a sampled walk takes one box");
    if ((size_t)nbox > ctx->walk_box_cap) {
        size_t cap = std::max<size_t>(64, (size_t)nbox * 2);
        RCHK(regrow(ctx, &ctx->walk_boxes, cap));
        RCHK(regrow(ctx, &ctx->walk_prefix, cap + 1));
        ctx->walk_box_cap = cap;
    }
    const law_tables* tab = boxes[0]->tab;
    std::vector<walk_box> hb((size_t)nbox);
    for (int b = 0; b < nbox; b++) {
        lagpm_box* box = boxes[b];
        if (box->tab != tab || box->p.precision != LAGPM_PREC_FP32)
            return lagpm_fail(ctx, LAGPM_E_INVAL, "boxes of one walk launch must share the FP32 force table");
        const double th = box->p.theta;
        walk_box& A = hb[(size_t)b];
        A.pp = box->pp; A.link = box->w_link; A.geo = box->w_geo; A.q0 = box->w_q0;
        A.q1 = box->w_q1; A.cellroot = box->w_cellroot; A.groups = box->groups; A.key = box->key;
        A.tvx = box->alt[0]; A.tvy = box->alt[1]; A.tvz = box->alt[2];  // free between sorts
        A.counters = box->counters; A.dsz = box->dsz;
        if (sample) {  // the sampled particles as groups of one (sizes: the sample's own words)
            A.groups = sample->groups;
            A.dsz = sample->dsz;
        }
        A.cell_shift = 3 * (LAGPM_KEY_BITS - box->level_cell);
        A.mcell = box->mcell;
        A.naccept = LAGPM_NACCEPT;
        // lengths in the units of pp (scaled by sc, see tree_prep_kernel)
        const double sc = lagpm_tree_scale(box);
        A.box = (float)(box->p.n * sc);
        A.rc2 = (float)(box->p.rsphere * box->p.rsphere * sc * sc);
        A.rcut = (float)(box->p.rsphere * sc);
        A.inv_theta2 = (float)(th > 0 ? 1.0 / (th * th) : 1e30);
        A.fact = (float)(facts[b] * box->p.mass / ((double)box->p.n * (double)box->p.n) / sc);
    }
    // pageable source: the copy is staged before the call returns, so hb may go out of scope
    HIPCHK(ctx, hipMemcpyAsync(ctx->walk_boxes, hb.data(), hb.size() * sizeof(walk_box), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(walk_prefix_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->walk_boxes, nbox, ctx->walk_prefix,
                       ctx->walk_ticket);
    HIPCHK(ctx, hipGetLastError());
    // each box's walk-start event, right before the launch (after the descriptor copy and the
    // prefix, so the phase time is the walk kernel's own)
    for (int b = 0; b < nbox; b++) HIPCHK(ctx, hipEventRecord(boxes[b]->ev[14], ctx->stream));
    // one workgroup per CU fills it (registers and LDS); reserve_cu CUs are left free for other
    // streams' kernels, in whole XCD rounds so every XCD keeps the same number of walk workgroups
    const int full = (int)(ctx->walk_stack_entries / ((size_t)STACK_MAX * WALK_WAVES));
    const int keep = full - ((std::max(0, reserve_cu) + WALK_XCDS - 1) / WALK_XCDS) * WALK_XCDS;
    const unsigned blocks = (unsigned)(keep >= WALK_XCDS ? keep : std::min(full, WALK_XCDS));
    auto kern = tab->logr2 ? (ctx->walk_census ? walk_f32_kernel<true, true> : walk_f32_kernel<true, false>)
                           : (ctx->walk_census ? walk_f32_kernel<false, true> : walk_f32_kernel<false, false>);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(WALK_THREADS), 0, ctx->stream, ctx->walk_boxes, ctx->walk_prefix, nbox,
                       tab->tab0, tab->tab12, ctx->walk_stack, ctx->walk_ticket);
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}

int lagpm_tree_walk_f64(lagpm_ctx* ctx, lagpm_box* box, double fact, const walk_sample* sample) {
    const double th = box->p.theta;
    const double cellw = (double)box->p.n / (double)box->mcell;
    walk64_args A;
    A.x = box->x; A.y = box->y; A.z = box->z; A.mass = box->p.mass; A.info = box->nd_info; A.child = box->nd_child;
    A.nd = box->nd_d; A.cellroot = box->cellroot; A.gbox = box->gbox;
    A.key = box->key; A.cell_shift = 3 * (LAGPM_KEY_BITS - box->level_cell); A.pgroup = box->pgroup; A.tab = box->tab->tabd;
    A.vx = box->vx; A.vy = box->vy; A.vz = box->vz; A.counters = box->counters; A.dsz = box->dsz; A.np = box->np;
    A.mcell = box->mcell; A.box = (double)box->p.n; A.cellw = cellw; A.rc2 = box->tab->rc2_eff; A.lam = box->tab->lam;
    A.rcut = box->p.rsphere; A.inv_theta2 = th > 0 ? 1.0 / (th * th) : 1e30;
    A.fact = fact * box->p.mass / ((double)box->p.n * (double)box->p.n);
    A.naccept = LAGPM_NACCEPT;
    A.sample = sample ? sample->groups : nullptr;
    A.out = sample ? sample->out : nullptr;
    A.nsample = sample ? sample->count : 0;
    if (box->np == 0) return LAGPM_OK;
    const size_t nthreads = sample ? sample->count : box->np;
    if (box->tab->logr2)
        hipLaunchKernelGGL(walk_f64_kernel<true>, dim3(blocks_for(nthreads, 128)), dim3(128), 0, box->st, A);
    else
        hipLaunchKernelGGL(walk_f64_kernel<false>, dim3(blocks_for(nthreads, 128)), dim3(128), 0, box->st, A);
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}

int lagpm_tree_readback_enqueue(lagpm_ctx* ctx, lagpm_box* box, tree_readback* dst) {
    HIPCHK(ctx, hipMemcpyAsync(dst->d, box->dsz, sizeof(dst->d), hipMemcpyDeviceToHost, box->st));
    HIPCHK(ctx, hipMemcpyAsync(dst->c, box->counters, sizeof(dst->c), hipMemcpyDeviceToHost, box->st));
    return LAGPM_OK;
}

int lagpm_tree_finish(lagpm_ctx* ctx, lagpm_box* box, bool* retry) {
    *retry = false;
    if (!ctx->readback1) HIPCHK(ctx, hipHostMalloc(&ctx->readback1, sizeof(tree_readback), hipHostMallocDefault));
    tree_readback* r = static_cast<tree_readback*>(ctx->readback1);
    RCHK(lagpm_tree_readback_enqueue(ctx, box, r));
    HIPCHK(ctx, hipStreamSynchronize(box->st));
    return lagpm_tree_finish_with(ctx, box, *r, retry);
}

int lagpm_tree_finish_with(lagpm_ctx* ctx, lagpm_box* box, const tree_readback& r, bool* retry) {
    *retry = false;
    const uint32_t* d = r.d;
    const unsigned long long* c = r.c;
    box->stats.n_pp = c[0];
    box->stats.n_pc = c[1];
    box->stats.n_pp_within = c[3];
    box->stats.n_nodes = d[DSZ_NODES];
    box->stats.n_groups = d[DSZ_GROUPS];
    if (d[DSZ_DEEPER]) {
        // the tree needs levels below the depth limit: redo it at full depth, with room for the
        // nodes the truncated build could not count
        box->depth = LAGPM_KEY_BITS;
        box->sorted = false;  // the keys are in order down to the old limit only: sort again
        lagpm_box_graphs_release(box);  // captured steps hold the old limit
        const size_t need = std::max<size_t>((size_t)d[DSZ_NODES], box->node_cap);
        if (d[DSZ_NODES] >= box->node_cap) RCHK(lagpm_tree_alloc(ctx, box, need + need / 2 + 4096));
        *retry = true;
        return LAGPM_E_OVERFLOW;
    }
    if (d[DSZ_OVERFLOW]) {
        // grow to the node count this input needs (groups follow from nodes) and rebuild
        size_t need = std::max<size_t>((size_t)d[DSZ_NODES], box->node_cap);
        RCHK(lagpm_tree_alloc(ctx, box, need + need / 4 + 4096));
        *retry = true;
        return LAGPM_E_OVERFLOW;
    }
    int deepest = box->level_cell;
    for (int l = box->level_cell; l <= LAGPM_KEY_BITS; l++)
        if (d[DSZ_LVL + l]) deepest = l;
    box->stats.deepest = deepest;
    if (box->depth >= LAGPM_KEY_BITS) {
        // first complete build at full depth: limit later builds to two levels below its deepest
        const int lim = std::min(LAGPM_KEY_BITS, deepest + 2);
        if (lim < box->depth) {
            box->depth = lim;
            lagpm_box_graphs_release(box);  // captured steps hold the old limit
        }
    }
    if (c[2]) return lagpm_fail(ctx, LAGPM_E_OVERFLOW, "tree walk stack overflow in " + std::to_string(c[2]) + " groups");
    return LAGPM_OK;
}

namespace {
// slot[ids[t]] = t: the sample index of each sampled caller index (slot starts at -1)
__global__ void sample_slot_kernel(const int64_t* __restrict__ ids, size_t count, int* __restrict__ slot) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) slot[ids[t]] = (int)t;
}
__global__ void fill_i32_kernel(int* __restrict__ a, size_t m, int v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
// every sample group starts as particle 0 alone (a valid group of one), then gets its particle
__global__ void sample_init_kernel(int2* __restrict__ groups, size_t count) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) groups[t] = make_int2(0, 1);
}
// sample t is the particle now at position p (the box's current order): a group of one at p
__global__ void sample_groups_kernel(const uint32_t* __restrict__ id, size_t np, const int* __restrict__ slot,
                                     int2* __restrict__ groups) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int t = slot[id[p]];
    if (t >= 0) groups[t] = make_int2((int)p, 1);
}
// the sample's size words: the box's (its overflow flag) with the group count replaced
__global__ void sample_dsz_kernel(const uint32_t* __restrict__ dsz, uint32_t* __restrict__ sdsz, uint32_t count) {
    for (int i = threadIdx.x; i < DSZ_COUNT; i += blockDim.x) sdsz[i] = i == DSZ_GROUPS ? count : dsz[i];
}
__global__ void sample_gather_kernel(const int2* __restrict__ groups, size_t count, const double* __restrict__ tx,
                                     const double* __restrict__ ty, const double* __restrict__ tz, double* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int p = groups[t].x;
    out[3 * t] = tx[p];
    out[3 * t + 1] = ty[p];
    out[3 * t + 2] = tz[p];
}
}  // namespace

// Walks the sampled particles of a built tree, each alone; acc (device, 3 per sample) gets their
// accelerations. ids (device) are caller indices, distinct and below np (checked by the caller).
int lagpm_tree_walk_sample(lagpm_ctx* ctx, lagpm_box* box, const int64_t* ids, size_t count, int* slot, int2* groups,
                           uint32_t* sdsz, double* acc) {
    const size_t np = box->np;
    hipStream_t st = box->st;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, slot, np, -1);
    hipLaunchKernelGGL(sample_slot_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, ids, count, slot);
    hipLaunchKernelGGL(sample_init_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, groups, count);
    hipLaunchKernelGGL(sample_groups_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, box->id, np, slot, groups);
    hipLaunchKernelGGL(sample_dsz_kernel, dim3(1), dim3(64), 0, st, box->dsz, sdsz, (uint32_t)count);
    HIPCHK(ctx, hipGetLastError());
    walk_sample smp{groups, sdsz, acc, count};
    if (box->p.precision == LAGPM_PREC_FP64) {
        RCHK(lagpm_tree_walk_f64(ctx, box, 1.0, &smp));
    } else {
        // the walk runs on the context stream and writes its kicks at the particles' positions
        if (st != ctx->stream) {
            HIPCHK(ctx, hipEventRecord(box->ev[10], st));
            HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, box->ev[10], 0));
        }
        const double one = 1.0;
        RCHK(lagpm_tree_walk_f32(ctx, &box, &one, 1, 0, &smp));
        hipLaunchKernelGGL(sample_gather_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, ctx->stream, groups, count,
                           (const double*)box->alt[0], (const double*)box->alt[1], (const double*)box->alt[2], acc);
        HIPCHK(ctx, hipGetLastError());
        if (st != ctx->stream) {
            HIPCHK(ctx, hipEventRecord(box->ev[11], ctx->stream));
            HIPCHK(ctx, hipStreamWaitEvent(st, box->ev[11], 0));
        }
    }
    return LAGPM_OK;
}

size_t lagpm_walk_stack_entries(int num_cu) {
    size_t blocks = (size_t)std::max(1, num_cu);  // one 16-wave workgroup per CU
    return blocks * WALK_WAVES * STACK_MAX;
}

// The octree of the last build in the box's own units: FP32 trees keep positions scaled by the
// tree scale s (see tree_prep_kernel) and masses in particle masses; the FP64 tree keeps cells.
extern "C" int lagpm_tree_nodes_get(lagpm_ctx* ctx, lagpm_box* box, size_t cap, size_t* nnodes, int32_t* info,
                                    double* com, double* quad, uint32_t* order) {
    if (!ctx || !box || !nnodes) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    *nnodes = 0;
    if (box->np == 0) return LAGPM_OK;
    if (!box->tree_built) return lagpm_fail(ctx, LAGPM_E_STATE, "no tree built: call lagpm_treecorrection first");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    uint32_t d[DSZ_COUNT];
    RCHK(lagpm_d2h(ctx, d, box->dsz, sizeof(d), box->st));
    if (d[DSZ_OVERFLOW]) return lagpm_fail(ctx, LAGPM_E_STATE, "the last build overflowed its node arrays");
    const size_t nn = std::min<size_t>(d[DSZ_NODES], box->node_cap);
    *nnodes = nn;
    if (cap < nn) return lagpm_fail(ctx, LAGPM_E_INVAL, "the arrays hold fewer nodes than the tree has (see *nnodes)");
    if (info) RCHK(lagpm_d2h(ctx, info, box->nd_info, nn * sizeof(int4), box->st));
    if (order) RCHK(lagpm_d2h(ctx, order, box->id, box->np * sizeof(uint32_t), box->st));
    if (!com && !quad) return LAGPM_OK;
    const double mass = box->p.mass;
    if (box->p.precision == LAGPM_PREC_FP64) {
        std::vector<double> t(12 * nn);
        RCHK(lagpm_d2h(ctx, t.data(), box->nd_d, t.size() * sizeof(double), box->st));
        for (size_t v = 0; v < nn; v++) {
            const double* a = &t[12 * v];
            if (com) { com[4 * v] = a[0]; com[4 * v + 1] = a[1]; com[4 * v + 2] = a[2]; com[4 * v + 3] = a[3] * mass; }
            if (quad) {
                for (int k = 0; k < 7; k++) quad[8 * v + k] = a[4 + k] * mass;
                quad[8 * v + 7] = a[11];
            }
        }
        return LAGPM_OK;
    }
    std::vector<float4> c(nn), q0(nn), q1(nn);
    RCHK(lagpm_d2h(ctx, c.data(), box->nd_com, nn * sizeof(float4), box->st));
    RCHK(lagpm_d2h(ctx, q0.data(), box->nd_q0, nn * sizeof(float4), box->st));
    RCHK(lagpm_d2h(ctx, q1.data(), box->nd_q1, nn * sizeof(float4), box->st));
    const double s = lagpm_tree_scale(box), s2 = s * s;
    for (size_t v = 0; v < nn; v++) {
        if (com) {
            com[4 * v] = c[v].x / s; com[4 * v + 1] = c[v].y / s; com[4 * v + 2] = c[v].z / s;
            com[4 * v + 3] = c[v].w * mass;
        }
        if (quad) {
            const double q[8] = {q0[v].x, q0[v].y, q0[v].z, q0[v].w, q1[v].x, q1[v].y, q1[v].z, q1[v].w};
            for (int k = 0; k < 7; k++) quad[8 * v + k] = q[k] * mass / s2;
            quad[8 * v + 7] = q[7] / s;
        }
    }
    return LAGPM_OK;
}
