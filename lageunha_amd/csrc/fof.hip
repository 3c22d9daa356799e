// fof.hip — PreFoF halo-candidate flags on the device (Cosmos/PreFoF.c:207-386, OST/fof.c:190-293).
//
// The reference bins particles into cells of width 4 (Cosmos/BasicCell.h:1), grows friends-of-
// friends groups inside each cell alone with a per-cell tree (pnew_fof_link), and keeps every
// group that lies within the link of a cell face or has at least FOF_MINNUM members
// (CountHaloCandidates). The groups are the connected components of the link graph, so here:
//   1. each particle gets a key (cell, sub-cell), the sub-cells no narrower than the link, and the
//      particles are radix-sorted by it (a cell's particles, and each sub-cell's, are contiguous);
//   2. one thread per particle searches the 27 neighbouring sub-cells of its own cell and joins
//      linked pairs in a lock-free union-find (roots are only ever hooked under smaller roots);
//   3. per group (at its root): member count and bounds, by atomics; then the flag rule.
// Pair distances and bounds are FP64 and use the reference's operations (relative coordinates
// X - 4*ix, sqrt(dx*dx + dy*dy + dz*dz) <= link, this file is built without FMA contraction), so
// the flags are the reference's bit for bit.
#include <vector>

#include "lagpm_internal.h"

namespace {

constexpr int FOF_SUB_MAX = 32;  // sub-cells per cell edge at most (15 key bits)
constexpr int FOF_SUB_BITS = 15;

__device__ inline uint32_t par_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void par_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving. Every parent is smaller than its child (hooks put the larger root
// under the smaller one; halving stores an ancestor), so the walk strictly decreases and ends.
__device__ inline uint32_t uf_root(uint32_t* par, uint32_t x) {
    while (true) {
        const uint32_t p = par_load(par + x);
        if (p == x) return x;
        const uint32_t g = par_load(par + p);
        if (g == p) return p;
        par_store(par + x, g);
        x = g;
    }
}
// Joins the sets of a and b. A failed hook means another thread hooked that root first (a
// successful operation elsewhere), so the retries are bounded by the number of hooks.
__device__ inline void uf_unite(uint32_t* par, uint32_t a, uint32_t b) {
    a = uf_root(par, a);
    b = uf_root(par, b);
    while (a != b) {
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(par + a, a, b);
        if (old == a) return;
        a = uf_root(par, old);
        b = uf_root(par, b);
    }
}

struct fof_grid {
    uint32_t mx;  // cells per edge: ceil(n / 4)
    int k;        // sub-cells per cell edge
    double kf;    // k / 4
};

__device__ inline uint32_t cell_of(double X, uint32_t mx) {
    const uint32_t c = (uint32_t)(X / 4.0);  // PreFoF.c:326-328 (X / 4 is exact)
    return c < mx ? c : mx - 1;
}

__global__ void fof_key_kernel(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                               size_t np, fof_grid G, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double P[3] = {x[i], y[i], z[i]};
    uint32_t c[3], s[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        c[a] = cell_of(P[a], G.mx);
        const double r = P[a] - 4.0 * (double)c[a];
        const int sb = (int)(r * G.kf);
        s[a] = (uint32_t)min(max(sb, 0), G.k - 1);
    }
    const uint64_t cell = (uint64_t)c[0] + (uint64_t)G.mx * ((uint64_t)c[1] + (uint64_t)G.mx * (uint64_t)c[2]);
    const uint64_t sub = (uint64_t)s[0] + (uint64_t)G.k * ((uint64_t)s[1] + (uint64_t)G.k * (uint64_t)s[2]);
    key[i] = (cell << FOF_SUB_BITS) | sub;
    val[i] = (uint32_t)i;
}

// positions relative to the cell corner (Dump2FoFPtl, PreFoF.c:38-54) in sorted order; set-up of
// the union-find and the per-group accumulators
__global__ void fof_gather_kernel(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                                  const uint32_t* __restrict__ val, size_t np, uint32_t mx, double* __restrict__ r,
                                  uint32_t* __restrict__ par, uint32_t* __restrict__ cnt, uint64_t* __restrict__ lo,
                                  uint64_t* __restrict__ hi) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t i = val[p];
    const double P[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r[3 * p + a] = P[a] - 4.0 * (double)cell_of(P[a], mx);
        lo[3 * p + a] = ~0ull;
        hi[3 * p + a] = 0ull;
    }
    par[p] = (uint32_t)p;
    cnt[p] = 0u;
}

__device__ inline uint32_t lower_bound_key(const uint64_t* __restrict__ K, uint32_t lo, uint32_t hi, uint64_t v) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (K[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// first sorted position of every cell c in [0, ncell] (start[ncell] = np)
__global__ void fof_cellstart_kernel(const uint64_t* __restrict__ key, uint32_t np, uint64_t ncell,
                                     uint32_t* __restrict__ start) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    start[c] = lower_bound_key(key, 0, np, c << FOF_SUB_BITS);
}

__global__ void fof_link_kernel(const uint64_t* __restrict__ key, const double* __restrict__ r,
                                const uint32_t* __restrict__ start, uint32_t np, int k, double link,
                                uint32_t* __restrict__ par) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint64_t kp = key[p];
    const uint64_t cell = kp >> FOF_SUB_BITS;
    const uint32_t sub = (uint32_t)(kp & ((1ull << FOF_SUB_BITS) - 1));
    const int sx = (int)(sub % (uint32_t)k), sy = (int)((sub / (uint32_t)k) % (uint32_t)k), sz = (int)(sub / (uint32_t)(k * k));
    const uint32_t cs = start[cell], ce = start[cell + 1];
    const double px = r[3 * (size_t)p], py = r[3 * (size_t)p + 1], pz = r[3 * (size_t)p + 2];
    // an ancestor of p (ancestors stay ancestors): a partner whose parent is it, or p itself, is
    // already in p's set, which saves the two root searches of a union (dense halo cells link
    // most of their pairs)
    uint32_t rp = uf_root(par, p);
    for (int dz = -1; dz <= 1; dz++) {
        const int tz = sz + dz;
        if (tz < 0 || tz >= k) continue;
        for (int dy = -1; dy <= 1; dy++) {
            const int ty = sy + dy;
            if (ty < 0 || ty >= k) continue;
            for (int dx = -1; dx <= 1; dx++) {
                const int tx = sx + dx;
                if (tx < 0 || tx >= k) continue;
                const uint32_t t = (uint32_t)(tx + k * (ty + k * tz));
                if (t < sub) continue;  // each unordered pair once: from the lower sub-cell
                const uint64_t tk = (cell << FOF_SUB_BITS) | t;
                uint32_t a = lower_bound_key(key, cs, ce, tk);
                const uint32_t b = lower_bound_key(key, a, ce, tk + 1);
                if (t == sub) a = p + 1;  // same sub-cell: the partners after p
                for (uint32_t q = a; q < b; q++) {
                    const double ex = fabs(px - r[3 * (size_t)q]), ey = fabs(py - r[3 * (size_t)q + 1]),
                                 ez = fabs(pz - r[3 * (size_t)q + 2]);
                    const double d = sqrt(ex * ex + ey * ey + ez * ez);  // OST/fof.c:272-276
                    if (d <= link) {
                        const uint32_t pq = par_load(par + q);
                        if (pq != rp && pq != p) {
                            uf_unite(par, p, q);
                            rp = uf_root(par, p);
                        }
                    }
                }
            }
        }
    }
}

__global__ void fof_stats_kernel(const double* __restrict__ r, uint32_t np, uint32_t* __restrict__ par,
                                 uint32_t* __restrict__ cnt, uint64_t* __restrict__ lo, uint64_t* __restrict__ hi) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < np;
    const uint32_t g = valid ? uf_root(par, p) : 0xffffffffu;
    // A wave whose particles all belong to one group (the sorted order puts a halo core's
    // particles together) reduces its count and bounds first and does one set of atomics: the
    // cores' roots otherwise take ~10^5 atomics each on one address.
    const uint32_t g0 = (uint32_t)__shfl((int)g, 0, 64);
    if (g0 != 0xffffffffu && __all(!valid || g == g0)) {
        const unsigned long long vm = __ballot(valid);
        double mn[3], mx[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double v = valid ? r[3 * (size_t)p + a] : 0.0;
            mn[a] = valid ? v : __longlong_as_double(0x7ff0000000000000ll);  // +inf
            mx[a] = valid ? v : -1.0;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                mn[a] = fmin(mn[a], __shfl_xor(mn[a], o, 64));
                mx[a] = fmax(mx[a], __shfl_xor(mx[a], o, 64));
            }
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(cnt + g0, (uint32_t)__popcll(vm));
#pragma unroll
            for (int a = 0; a < 3; a++) {
                atomicMin((unsigned long long*)(lo + 3 * (size_t)g0 + a), (unsigned long long)__double_as_longlong(mn[a]));
                atomicMax((unsigned long long*)(hi + 3 * (size_t)g0 + a), (unsigned long long)__double_as_longlong(mx[a]));
            }
        }
        return;
    }
    if (!valid) return;
    atomicAdd(cnt + g, 1u);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // relative coordinates are >= +0, so their bit patterns order like their values
        const uint64_t v = (uint64_t)__double_as_longlong(r[3 * (size_t)p + a]);
        atomicMin((unsigned long long*)(lo + 3 * (size_t)g + a), (unsigned long long)v);
        atomicMax((unsigned long long*)(hi + 3 * (size_t)g + a), (unsigned long long)v);
    }
}

// CountHaloCandidates, PreFoF.c:238-251: a group inside [link, 4 - link) on every axis is kept
// only with minnum members or more; any other group is kept
__global__ void fof_flag_kernel(const uint32_t* __restrict__ val, const uint32_t* __restrict__ id, uint32_t np,
                                uint32_t* __restrict__ par, const uint32_t* __restrict__ cnt,
                                const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi, double link,
                                uint32_t minnum, uint8_t* __restrict__ flag) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t g = uf_root(par, p);
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double l = __longlong_as_double((long long)lo[3 * (size_t)g + a]);
        const double h = __longlong_as_double((long long)hi[3 * (size_t)g + a]);
        inside = inside && l > link && h < 4.0 - link;
    }
    flag[id[val[p]]] = (uint8_t)(inside ? (cnt[g] >= minnum) : 1);
}

unsigned grid_of(size_t n, int t) { return (unsigned)std::max<size_t>(1, (n + t - 1) / t); }

struct dev_buf {
    void* p = nullptr;
    ~dev_buf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int lagpm_prefof(lagpm_ctx* ctx, lagpm_box* box, float fof_link, int minnum, uint8_t* flags,
                            size_t* nflagged) {
    if (!ctx || !box) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    if (!box->uploaded) return lagpm_fail(ctx, LAGPM_E_STATE, "no particles uploaded");
    if (!(fof_link > 0.0f) || !(fof_link < 1e30f)) return lagpm_fail(ctx, LAGPM_E_INVAL, "fof_link must be positive and finite");
    if (minnum < 1) return lagpm_fail(ctx, LAGPM_E_INVAL, "minnum must be at least 1");
    const size_t np = box->np;
    if (nflagged) *nflagged = 0;
    if (np == 0) return LAGPM_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const double link = (double)fof_link;
    fof_grid G;
    G.mx = (uint32_t)((box->p.n + 3) / 4);
    // sub-cells strictly wider than the link, so a linked pair is in the same or adjacent sub-cells
    int k = (int)std::floor(4.0 / link);
    k = std::max(1, std::min(k, FOF_SUB_MAX));
    while (k > 1 && 4.0 / k - link < 1e-9 * 4.0) k--;
    G.k = k;
    G.kf = (double)k / 4.0;
    const uint64_t ncell = (uint64_t)G.mx * G.mx * G.mx;
    int cbits = 1;
    while ((1ull << cbits) < ncell + 1) cbits++;
    const int bits = cbits + FOF_SUB_BITS;
    hipStream_t st = box->st;
    // buffers: keys and values (double-buffered for the sort), relative positions, union-find,
    // group accumulators, cell starts, sort scratch, flags
    const size_t sort_tmp = lagpm_sort_tmp_bytes(np);
    const size_t off_key = 0, off_key2 = off_key + 8 * np, off_r = off_key2 + 8 * np, off_lo = off_r + 24 * np,
                 off_hi = off_lo + 24 * np, off_val = off_hi + 24 * np, off_val2 = off_val + 4 * np,
                 off_par = off_val2 + 4 * np, off_cnt = off_par + 4 * np, off_start = off_cnt + 4 * np,
                 off_tmp = (off_start + 4 * (ncell + 1) + 255) / 256 * 256, off_flag = off_tmp + (sort_tmp + 255) / 256 * 256,
                 total = off_flag + np;
    dev_buf mem;
    if (hipMalloc(&mem.p, total) != hipSuccess) {
        (void)hipGetLastError();
        return lagpm_fail(ctx, LAGPM_E_NOMEM, "PreFoF scratch (" + std::to_string(total) + " bytes) does not fit");
    }
    char* base = (char*)mem.p;
    uint64_t* key = (uint64_t*)(base + off_key);
    uint64_t* key2 = (uint64_t*)(base + off_key2);
    double* r = (double*)(base + off_r);
    uint64_t* lo = (uint64_t*)(base + off_lo);
    uint64_t* hi = (uint64_t*)(base + off_hi);
    uint32_t* val = (uint32_t*)(base + off_val);
    uint32_t* val2 = (uint32_t*)(base + off_val2);
    uint32_t* par = (uint32_t*)(base + off_par);
    uint32_t* cnt = (uint32_t*)(base + off_cnt);
    uint32_t* start = (uint32_t*)(base + off_start);
    uint32_t* tmp = (uint32_t*)(base + off_tmp);
    uint8_t* dflag = (uint8_t*)(base + off_flag);
    hipLaunchKernelGGL(fof_key_kernel, dim3(grid_of(np, 256)), dim3(256), 0, st, box->x, box->y, box->z, np, G, key, val);
    HIPCHK(ctx, hipGetLastError());
    bool in_alt = false;
    RCHK(lagpm_radix_sort_pairs(ctx, st, tmp, key, key2, val, val2, np, bits, &in_alt));
    if (in_alt) {
        std::swap(key, key2);
        std::swap(val, val2);
    }
    hipLaunchKernelGGL(fof_gather_kernel, dim3(grid_of(np, 256)), dim3(256), 0, st, box->x, box->y, box->z, val, np, G.mx,
                       r, par, cnt, lo, hi);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(fof_cellstart_kernel, dim3(grid_of(ncell + 1, 256)), dim3(256), 0, st, key, (uint32_t)np, ncell,
                       start);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(fof_link_kernel, dim3(grid_of(np, 256)), dim3(256), 0, st, key, r, start, (uint32_t)np, G.k, link,
                       par);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(fof_stats_kernel, dim3(grid_of(np, 256)), dim3(256), 0, st, r, (uint32_t)np, par, cnt, lo, hi);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(fof_flag_kernel, dim3(grid_of(np, 256)), dim3(256), 0, st, val, box->id, (uint32_t)np, par, cnt,
                       lo, hi, link, (uint32_t)minnum, dflag);
    HIPCHK(ctx, hipGetLastError());
    std::vector<uint8_t> h(np);
    RCHK(lagpm_d2h(ctx, h.data(), dflag, np, st));
    size_t c = 0;
    for (size_t i = 0; i < np; i++) c += h[i];
    if (flags) std::copy(h.begin(), h.end(), flags);
    if (nflagged) *nflagged = c;
    return LAGPM_OK;
}
