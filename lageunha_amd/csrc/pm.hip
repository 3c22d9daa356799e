// pm.hip — PM long-range force: TSC deposit, rocFFT Poisson solve with the Gaussian-filtered
// Green function, FDA4 gradient mesh, TSC interpolation and kick.
//
// Reference semantics (paths relative to the reference tree):
//   deposit       Cosmos/pmmain.h:3-78 (Get_TSC_Den) + periodic fold RkS/gmigrate.c:212-275
//   mean shift    Cosmos/pmmain.c:304 (grid initialised to -1 before the fold)
//   Green         Cosmos/psolver.c:13-86
//   gradient      Cosmos/fda4.c:46-47, Cosmos/fda4.h:302-384 (FDA4_SLOW)
//   interpolation Cosmos/fda4.h:11-102 (FDA4_HEAD weights), :276-300 (FDA4_TAIL sum + kick)
//
// This file is compiled with -ffp-contract=off: every float statement rounds once, exactly
// as the reference's C statements do, so the FP32 path is bit-comparable with the oracle.
//
// Deposit design: particles arrive in Morton order; a block takes 1024 consecutive particles,
// reduces their TSC footprint to a bounding box and, when it fits 4096 cells, accumulates the
// 27 contributions per particle into LDS with 64-bit integer atomics before flushing non-zero
// cells with one global 64-bit atomic each. Contributions are fixed point (x 2^S, round to
// nearest even), so the sum is exact and independent of the order: deterministic.
#include "lagpm_internal.h"
#include "wave_ops.h"

#include <cmath>
#include <type_traits>

namespace {

#ifndef LAGPM_DEP_PPT
#define LAGPM_DEP_PPT 4
#endif
// Resident deposit blocks per CU (the LDS box allows 4). Two leave room for the tree build's
// kernels, which run beside the deposit in a step: 4 per CU deposit faster alone but hold the CUs'
// LDS while the build waits, and the whole step is slower.
#ifndef LAGPM_DEP_BLOCKS_PER_CU
#define LAGPM_DEP_BLOCKS_PER_CU 2
#endif
#ifndef LAGPM_DEP_THREADS
#define LAGPM_DEP_THREADS 256
#endif
constexpr int DEP_THREADS = LAGPM_DEP_THREADS;
constexpr int DEP_PPT = LAGPM_DEP_PPT;  // particles per thread and chunk
constexpr int DEP_BLOCKS_PER_CU = LAGPM_DEP_BLOCKS_PER_CU;
constexpr int DEP_CHUNK = DEP_THREADS * DEP_PPT;
constexpr int DEP_BOX = 16;  // the LDS box of a chunk: 16^3 cells, index x + 16 (y + 16 z)
constexpr int DEP_LDS = DEP_BOX * DEP_BOX * DEP_BOX;

// periodic wrap of a cell index within one period of the box ([-n, 2n)): compares, no division
__device__ inline int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Deposit weights, x axis carries the mass (pmmain.h:11-40).
__device__ inline void tsc_axis_mass(double xp, float pmas0, int& near, int& s, float w[3]) {
    double r = rint(xp);
    near = (int)r;
    double xmin = xp - r;
    s = signbit(xmin) ? -1 : 1;
    float xd1 = (float)fabs(xmin);
    float p05 = (float)((double)pmas0 * 0.5);
    float w1 = (float)((0.75 - (double)(xd1 * xd1)) * (double)pmas0);
    float w3 = (float)((double)p05 * (0.25 + (double)xd1 * ((double)xd1 - 1.0)));
    float w2 = w3 + pmas0 * xd1;
    w[0] = w1; w[1] = w2; w[2] = w3;
}
__device__ inline void tsc_axis_unit(double xp, int& near, int& s, float w[3]) {
    double r = rint(xp);
    near = (int)r;
    double xmin = xp - r;
    s = signbit(xmin) ? -1 : 1;
    float xd1 = (float)fabs(xmin);
    float w1 = (float)(0.75 - (double)(xd1 * xd1));
    float w3 = (float)(0.5 * (0.25 + (double)xd1 * ((double)xd1 - 1.0)));
    float w2 = w3 + xd1;
    w[0] = w1; w[1] = w2; w[2] = w3;
}
// FP64 weights for the plain FP64 path
__device__ inline void tsc_axis_d(double xp, int& near, int& s, double w[3]) {
    double r = rint(xp);
    near = (int)r;
    double xmin = xp - r;
    s = signbit(xmin) ? -1 : 1;
    double d = fabs(xmin);
    w[0] = 0.75 - d * d;
    w[2] = 0.5 * (0.25 + d * (d - 1.0));
    w[1] = w[2] + d;
}

__global__ void zero_u64_kernel(unsigned long long* __restrict__ a, size_t m) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) a[i] = 0ull;
}

// Deposit chunks: runs of at most DEP_CHUNK consecutive (Morton-sorted) particles that stay
// inside one Morton block of <= 13 mesh cells per side, so every chunk's TSC footprint fits the
// 16^3 LDS box. Block b's particles are the sorted keys with block prefix b; the block is cut into
// ceil(count / DEP_CHUNK) chunks. The list is built per block (a binary search for each block's
// first particle, a scan over the blocks' chunk counts, the starts written per block), so no pass
// touches every particle.
//
// bstart[b] = first particle of block b (b <= nblk; bstart[nblk] = N), bcnt[b] = its chunk count;
// the accumulator is cleared (m cells) in the same launch.
__global__ void dep_blocks_kernel(const uint64_t* __restrict__ K, size_t N, int sh, uint32_t nblk,
                                  uint32_t* __restrict__ bstart, uint32_t* __restrict__ bcnt,
                                  unsigned long long* __restrict__ acc, size_t m) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = t0; i < m; i += stride) acc[i] = 0ull;
    auto first_at_least = [&](uint64_t pfx) -> size_t {  // first index whose prefix is >= pfx
        size_t lo = 0, hi = N;
        while (lo < hi) {
            const size_t mid = lo + ((hi - lo) >> 1);
            if ((K[mid] >> sh) >= pfx) hi = mid;
            else lo = mid + 1;
        }
        return lo;
    };
    for (size_t b = t0; b <= nblk; b += stride) {
        const size_t s = b == nblk ? N : first_at_least(b);
        bstart[b] = (uint32_t)s;
        if (b < nblk) {
            const size_t e = b + 1 == nblk ? N : first_at_least(b + 1);
            bcnt[b] = (uint32_t)((e - s + DEP_CHUNK - 1) / DEP_CHUNK);
        }
    }
}
// chunk starts of block b from its first particle and chunk offset; the sentinel start[nchunk] = N
__global__ void dep_chunks_kernel(const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bcnt,
                                  const uint32_t* __restrict__ boff, uint32_t nblk, size_t N,
                                  uint32_t* __restrict__ start, const uint32_t* __restrict__ nchunk) {
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) start[*nchunk] = (uint32_t)N;
    if (b >= nblk) return;
    const uint32_t c = bcnt[b], o = boff[b], s = bstart[b];
    for (uint32_t j = 0; j < c; j++) start[o + j] = s + j * (uint32_t)DEP_CHUNK;
}

// LDS box cell of offsets (x, y, z) in [0, 16)^3 from the chunk's corner. Rows of 16 cells, with
// bits 2-3 of x flipped by y mod 4: the 64 lanes of a wave hold 64 consecutive Morton-ordered
// particles, on a lattice a 4x4x4 block of cells, and each atomic of the 27 then spreads them over
// 16 bank pairs (4 lanes each, the minimum for 64 eight-byte operands) instead of 4.
__device__ inline int dep_cell(int x, int y, int z) {
    return (x ^ ((y & 3) << 2)) + DEP_BOX * y + DEP_BOX * DEP_BOX * z;
}

// Fixed point of a contribution 0 <= c * 2^s < 2^52: c * 2^s is exact (a power-of-two scale), and
// adding 2^52 in the same fused operation rounds it once to an integer, half to even (llrint);
// that integer is the low 52 bits of the sum's pattern. The host checks the bound; above it the
// general conversion is taken.
template <bool MAGIC>
__device__ inline unsigned long long dep_fixed(double c, double scale) {
    if constexpr (MAGIC) {
        return (unsigned long long)__double_as_longlong(fma(c, scale, 0x1p52)) & 0x000fffffffffffffull;
    } else {
        return (unsigned long long)__double2ll_rn(c * scale);
    }
}

// Persistent blocks over the chunk list. Thread t takes the particles t, t + DEP_THREADS, ... of a
// chunk (coalesced loads; a wave holds 64 consecutive particles). Per chunk: the box of nearest
// cells from the positions already in registers, the loads of the next chunk's positions issued,
// then the 27 contributions of each particle added to the 16^3 LDS box (or straight to the mesh
// when the box does not fit), and the box added to the mesh once per nonzero cell, which is
// cleared in the same pass for the next chunk. Every sum is of integers, so the result is exact
// and independent of the order of the additions.
template <bool F64, bool MAGIC>
__global__ void __launch_bounds__(DEP_THREADS, DEP_BLOCKS_PER_CU) deposit_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                                              const double* __restrict__ pz, const uint32_t* __restrict__ start,
                                                              const uint32_t* __restrict__ nchunk, unsigned chunk_cap,
                                                              size_t npart, int n, float mass, double scale,
                                                              unsigned long long* __restrict__ acc) {
    __shared__ unsigned long long box[DEP_LDS];
    __shared__ int red[6][DEP_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const unsigned nch = min(*nchunk, chunk_cap);
    // XCD x (block b runs on XCD b % 8) walks the x-th eighth of the chunks, so neighbouring
    // chunks, which add to neighbouring mesh cells, share one L2
    const unsigned per = gridDim.x / 8u, x = blockIdx.x % 8u, j = blockIdx.x / 8u;
    const unsigned clo = (unsigned)(((unsigned long long)nch * x) / 8u), chi = (unsigned)(((unsigned long long)nch * (x + 1)) / 8u);
    unsigned c = clo + j;
    if (c >= chi) return;
    for (int q = tid; q < DEP_LDS; q += DEP_THREADS) box[q] = 0ull;

    // this chunk's particles (clamped to the particle count) and their positions
    auto bounds = [&](unsigned ch, size_t& b, size_t& e) {
        e = min((size_t)start[ch + 1], npart);
        b = min((size_t)start[ch], e);
    };
    size_t base, end;
    bounds(c, base, end);
    double X[DEP_PPT], Y[DEP_PPT], Z[DEP_PPT];
#pragma unroll
    for (int k = 0; k < DEP_PPT; k++) {
        const size_t i = base + (size_t)k * DEP_THREADS + tid;
        const bool in = i < end;
        X[k] = in ? px[i] : 0.0; Y[k] = in ? py[i] : 0.0; Z[k] = in ? pz[i] : 0.0;
    }
    for (;;) {
        int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
        for (int k = 0; k < DEP_PPT; k++) {
            if (base + (size_t)k * DEP_THREADS + tid < end) {
                const int nx = (int)rint(X[k]), ny = (int)rint(Y[k]), nz = (int)rint(Z[k]);
                lo[0] = min(lo[0], nx); hi[0] = max(hi[0], nx);
                lo[1] = min(lo[1], ny); hi[1] = max(hi[1], ny);
                lo[2] = min(lo[2], nz); hi[2] = max(hi[2], nz);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            for (int o = 32; o > 0; o >>= 1) {
                lo[a] = min(lo[a], __shfl_xor(lo[a], o, 64));
                hi[a] = max(hi[a], __shfl_xor(hi[a], o, 64));
            }
            if (lane == 0) { red[a][wid] = lo[a]; red[3 + a][wid] = hi[a]; }
        }
        // the next chunk of this block and the loads of its positions, in flight during this one
        const unsigned cn = c + per;
        const bool more = cn < chi;
        size_t nbase = 0, nend = 0;
        if (more) bounds(cn, nbase, nend);
        double NX[DEP_PPT], NY[DEP_PPT], NZ[DEP_PPT];
#pragma unroll
        for (int k = 0; k < DEP_PPT; k++) {
            const size_t i = nbase + (size_t)k * DEP_THREADS + tid;
            const bool in = i < nend;
            NX[k] = in ? px[i] : 0.0; NY[k] = in ? py[i] : 0.0; NZ[k] = in ? pz[i] : 0.0;
        }
        __syncthreads();  // red written; the box cleared by the previous pass
        int blo[3], bdim[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            int l = INT_MAX, h = INT_MIN;
            for (int w = 0; w < DEP_THREADS / 64; w++) { l = min(l, red[a][w]); h = max(h, red[3 + a][w]); }
            blo[a] = l - 1;
            bdim[a] = (h + 1) - (l - 1) + 1;
        }
        // (near lies in [0, n] for wrapped positions, so every index below is in [-1, n + 1])
        const bool use_lds = bdim[0] > 0 && bdim[0] <= DEP_BOX && bdim[1] <= DEP_BOX && bdim[2] <= DEP_BOX;
#pragma unroll
        for (int k = 0; k < DEP_PPT; k++) {
            if (base + (size_t)k * DEP_THREADS + tid >= end) continue;
            int ix, iy, iz, sx, sy, sz;
            float wx[3], wy[3], wz[3];
            double dwx[3], dwy[3], dwz[3];
            if constexpr (F64) {
                tsc_axis_d(X[k], ix, sx, dwx);
                tsc_axis_d(Y[k], iy, sy, dwy);
                tsc_axis_d(Z[k], iz, sz, dwz);
            } else {
                tsc_axis_mass(X[k], mass, ix, sx, wx);
                tsc_axis_unit(Y[k], iy, sy, wy);
                tsc_axis_unit(Z[k], iz, sz, wz);
            }
            const int xi[3] = {ix, ix + sx, ix - sx};
            const int yi[3] = {iy, iy + sy, iy - sy};
            const int zi[3] = {iz, iz + sz, iz - sz};
            // the fixed-point contribution of node (a, b, c): the reference's float product
            // (wx wy) wz (pmmain.h:11-40), or the FP64 product for the FP64 path
            auto contrib = [&](int a, int b, int cc) -> unsigned long long {
                if constexpr (F64) {
                    return dep_fixed<MAGIC>((double)mass * dwx[a] * dwy[b] * dwz[cc], scale);
                } else {
                    const float wxy = wx[a] * wy[b];
                    return dep_fixed<MAGIC>((double)(wxy * wz[cc]), scale);
                }
            };
            if (use_lds) {
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
#pragma unroll
                    for (int b = 0; b < 3; b++)
#pragma unroll
                        for (int a = 0; a < 3; a++)
                            atomicAdd(&box[dep_cell(xi[a] - blo[0], yi[b] - blo[1], zi[cc] - blo[2])], contrib(a, b, cc));
            } else {
                size_t gx[3], gy[3], gz[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    gx[a] = (size_t)wrapc(xi[a], n);
                    gy[a] = (size_t)n * (size_t)wrapc(yi[a], n);
                    gz[a] = (size_t)n * (size_t)n * (size_t)wrapc(zi[a], n);
                }
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
#pragma unroll
                    for (int b = 0; b < 3; b++)
#pragma unroll
                        for (int a = 0; a < 3; a++) atomicAdd(&acc[gx[a] + gy[b] + gz[cc]], contrib(a, b, cc));
            }
        }
        __syncthreads();  // the box is complete
        if (use_lds) {
            for (int q = tid; q < DEP_LDS; q += DEP_THREADS) {
                const unsigned long long v = box[q];
                if (v == 0ull) continue;  // cells outside the chunk's box are never written
                box[q] = 0ull;
                const int d = q / (DEP_BOX * DEP_BOX), b = (q / DEP_BOX) & (DEP_BOX - 1);
                const int a = (q & (DEP_BOX - 1)) ^ ((b & 3) << 2);
                const size_t g = (size_t)wrapc(blo[0] + a, n) + (size_t)n * ((size_t)wrapc(blo[1] + b, n) + (size_t)n * wrapc(blo[2] + d, n));
                atomicAdd(&acc[g], v);
            }
        }
        if (!more) break;
        c = cn; base = nbase; end = nend;
#pragma unroll
        for (int k = 0; k < DEP_PPT; k++) { X[k] = NX[k]; Y[k] = NY[k]; Z[k] = NZ[k]; }
    }
}

// The FFT mesh is the in-place real layout: rows of n values padded to mx = 2(n/2+1), the
// layout of the reference's FFTW grid (Cosmos/pmmain.c:328-334). It holds the density, then
// the spectrum, then the potential.
template <typename T>
__global__ void convert_kernel(const unsigned long long* __restrict__ acc, int n, double inv_scale, double mean,
                               T* __restrict__ rho) {
    const size_t m = (size_t)n * n * n, mx = 2 * (size_t)(n / 2 + 1);
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < m; i += stride) rho[i % n + mx * (i / n)] = (T)((double)(long long)acc[i] * inv_scale - mean);
}

__device__ inline double tsc_w2(int n, double m) {
    if (m == 0.0) return 1.0;
    double h = M_PI * m / (double)n;
    double s = sin(h) / h;
    double s3 = s * s * s;
    return s3 * s3;
}

// Green factor of psolver.c:59-72 for wave numbers (i, wy, wz), in the reference's own operations:
// the squared wave number in float (exact for these integers), the factor in double.
__device__ inline double green_factor(int n, int i, float wy, float wz, double sfact, int deconvolve) {
    const double wave = (double)((float)(i * i) + wy * wy + wz * wz);
    double factor = ((-1.0 / M_PI) / wave) * exp(wave * sfact);
    if (deconvolve) factor /= tsc_w2(n, (double)i) * tsc_w2(n, (double)wy) * tsc_w2(n, (double)wz);
    return factor;
}

// Green multiply over the half-complex spectrum, layout (i + h*(j + n*k)), psolver.c:49-76.
// One workgroup per segment of `segs` of one half-complex row (fixed j, k): the row's indices and
// its y and z wave numbers are workgroup-uniform, so no cell divides its linear index.
template <typename T>
__global__ void green_kernel(T* __restrict__ spec, int n, int segs, double sfact, double irngc, int deconvolve) {
    const int h = n / 2 + 1;
    const unsigned row = blockIdx.x / (unsigned)segs;  // row = j + n k
    const int i = (int)(blockIdx.x - row * (unsigned)segs) * (int)blockDim.x + (int)threadIdx.x;
    if (i >= h) return;
    const int j = (int)(row % (unsigned)n), k = (int)(row / (unsigned)n);
    const size_t q = (size_t)i + (size_t)h * row;
    const double rn = n;
    float wy = (float)j;
    if (wy > rn / 2) wy = wy - (float)rn;
    float wz = (float)k;
    if (wz > rn / 2) wz = wz - (float)rn;
    if (i == 0 && j == 0 && k == 0) {
        spec[2 * q] = (T)0;
        spec[2 * q + 1] = (T)0;
        return;
    }
    const double factor = green_factor(n, i, wy, wz, sfact, deconvolve);
    spec[2 * q] = (T)(factor * (double)spec[2 * q] * irngc);
    spec[2 * q + 1] = (T)(factor * (double)spec[2 * q + 1] * irngc);
}

// The factor depends on i, |wy| and |wz| only (wave numbers enter squared, and the TSC window is
// even: sin(-x)/(-x) = sin(x)/x in floating point, as sin is odd), so a context keeps it once per
// (n, efold, deconvolution) in a table of (n/2+1)^3 doubles, index i + h (|wy| + h |wz|), and
// the step's multiply streams the spectrum against it: the same double factor, the same products.
__global__ void green_table_kernel(double* __restrict__ tab, int n, double sfact, int deconvolve) {
    const int h = n / 2 + 1;
    const size_t cnt = (size_t)h * h * h;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % h);
        const size_t r = q / h;
        const int ay = (int)(r % h), az = (int)(r / h);
        tab[q] = (i | ay | az) ? green_factor(n, i, (float)ay, (float)az, sfact, deconvolve) : 0.0;
    }
}
template <typename T>
__global__ void green_tab_kernel(T* __restrict__ spec, const double* __restrict__ tab, int n, int segs, double irngc) {
    const int h = n / 2 + 1;
    const unsigned row = blockIdx.x / (unsigned)segs;  // row = j + n k
    const int i = (int)(blockIdx.x - row * (unsigned)segs) * (int)blockDim.x + (int)threadIdx.x;
    if (i >= h) return;
    const int j = (int)(row % (unsigned)n), k = (int)(row / (unsigned)n);
    const int ay = j <= n / 2 ? j : n - j, az = k <= n / 2 ? k : n - k;
    const size_t q = (size_t)i + (size_t)h * row;
    if (i == 0 && j == 0 && k == 0) {
        spec[2 * q] = (T)0;
        spec[2 * q + 1] = (T)0;
        return;
    }
    const double factor = tab[(size_t)i + (size_t)h * ((size_t)ay + (size_t)h * az)];
    spec[2 * q] = (T)(factor * (double)spec[2 * q] * irngc);
    spec[2 * q + 1] = (T)(factor * (double)spec[2 * q + 1] * irngc);
}

__device__ inline float fda4_f(float m1, float p1, float m2, float p2) {
    const float ffact1 = (float)((4. / 3.) / 2.);
    const float ffact2 = (float)((-1. / 3.) / 4.);
    float a = ffact1 * (m1 - p1);
    float b = ffact2 * (m2 - p2);
    return a + b;
}
__device__ inline double fda4_d(double m1, double p1, double m2, double p2) {
    return (2.0 / 3.0) * (m1 - p1) + (-1.0 / 12.0) * (m2 - p2);
}

// FDA4 gradient mesh: one thread per cell, 12 neighbour reads (cached), one 16/32-byte store.
// A workgroup takes one segment of `segs` of one mesh row (fixed j, k), so the row's j and k are
// workgroup-uniform and every index is an add; neighbours within +-2 wrap with one compare. Rows
// go to the XCDs in contiguous ranges (xcd_block): the +-2 planes and rows a row reads are then
// mostly in the same L2.
__device__ inline int wrap2(int a, int n) {  // a in [-n, 2n)
    return a < 0 ? a + n : (a >= n ? a - n : a);
}
template <typename T, typename T4>
__global__ void gradient_kernel(const T* __restrict__ phi, int n, int segs, T4* __restrict__ grad) {
    const unsigned lb = lagpm_wave::xcd_block(blockIdx.x, gridDim.x);
    const unsigned row = lb / (unsigned)segs;  // row = j + n k
    const int i = (int)(lb - row * (unsigned)segs) * (int)blockDim.x + (int)threadIdx.x;
    if (i >= n) return;
    const int j = (int)(row % (unsigned)n), k = (int)(row / (unsigned)n);
    const size_t mx = 2 * (size_t)(n / 2 + 1), plane = mx * n;  // padded rows of the FFT mesh
    const size_t rowj = mx * j + plane * k;
    auto X = [&](int a) { return phi[(size_t)wrap2(a, n) + rowj]; };
    auto Y = [&](int b) { return phi[(size_t)i + mx * wrap2(b, n) + plane * k]; };
    auto Z = [&](int c) { return phi[(size_t)i + mx * j + plane * wrap2(c, n)]; };
    T4 g;
    if constexpr (sizeof(T) == 4) {
        g.x = fda4_f(X(i - 1), X(i + 1), X(i - 2), X(i + 2));
        g.y = fda4_f(Y(j - 1), Y(j + 1), Y(j - 2), Y(j + 2));
        g.z = fda4_f(Z(k - 1), Z(k + 1), Z(k - 2), Z(k + 2));
    } else {
        g.x = fda4_d(X(i - 1), X(i + 1), X(i - 2), X(i + 2));
        g.y = fda4_d(Y(j - 1), Y(j + 1), Y(j - 2), Y(j + 2));
        g.z = fda4_d(Z(k - 1), Z(k + 1), Z(k - 2), Z(k + 2));
    }
    g.w = phi[(size_t)i + rowj];
    grad[(size_t)i + (size_t)n * row] = g;
}

// Interpolation weights of FDA4_HEAD: xmin is a float there (fda4.h:40-66).
__device__ inline void fda4_axis(double xp, int& near, int& s, float w[3]) {
    double r = rint(xp);
    near = (int)r;
    float xmin = (float)(xp - r);
    s = signbit(xmin) ? -1 : 1;
    float xd1 = fabsf(xmin);
    float w1 = (float)(0.75 - (double)(xd1 * xd1));
    float w3 = (float)(0.5 * (0.25 + (double)xd1 * ((double)xd1 - 1.0)));
    float w2 = w3 + xd1;
    w[0] = w1; w[1] = w2; w[2] = w3;
}

// TSC interpolation of the FDA4 gradient mesh and the kick (fda4.c:31-153). The 256 particles of
// a block are consecutive in Morton order, so their 27-cell stencils usually cover a small box of
// cells: that box is staged in LDS once (coalesced rows, periodic indices) and the 27 reads per
// particle come from LDS instead of the caches. A block whose box does not fit (its particles
// straddle a periodic face or two distant Morton cells) reads the mesh directly. Same cells, same
// weights and the same order of additions either way, so the kick is the same bit for bit.
constexpr int INTERP_THREADS = 256;
constexpr int INTERP_LDS = 1536;  // float4 cells of one block's box (24 KB)

__device__ inline void interp_sum(const float wx[3], const float wy[3], const float wz[3], const int xi[3],
                                  const int yi[3], const int zi[3], size_t sy, size_t sz, const float4* __restrict__ g,
                                  float& fx, float& fy, float& fz) {
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const size_t row = sy * (size_t)yi[b] + sz * (size_t)zi[c];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float wxy = wx[a] * wy[b];
                float wt = wxy * wz[c];
                float4 v = g[row + xi[a]];
                if (a == 0 && b == 0 && c == 0) {
                    fx = wt * v.x; fy = wt * v.y; fz = wt * v.z;
                } else {
                    fx = fx + wt * v.x; fy = fy + wt * v.y; fz = fz + wt * v.z;
                }
            }
        }
}

__global__ void __launch_bounds__(INTERP_THREADS) interp_kick_f32(const double* __restrict__ px, const double* __restrict__ py,
                                const double* __restrict__ pz, double* __restrict__ vx, double* __restrict__ vy,
                                double* __restrict__ vz, size_t np, int n, const float4* __restrict__ grad,
                                const double* __restrict__ fac) {
    __shared__ float4 sg[INTERP_LDS];
    __shared__ int wred[6][INTERP_THREADS / 64];
    __shared__ int box[6];
    const float fact2 = (float)fac[1];
    const size_t p = (size_t)lagpm_wave::xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = p < np;
    int near[3] = {0, 0, 0}, sgn[3] = {1, 1, 1};
    float wx[3] = {0.f, 0.f, 0.f}, wy[3] = {0.f, 0.f, 0.f}, wz[3] = {0.f, 0.f, 0.f};
    if (valid) {
        fda4_axis(px[p], near[0], sgn[0], wx);
        fda4_axis(py[p], near[1], sgn[1], wy);
        fda4_axis(pz[p], near[2], sgn[2], wz);
    }
    // the block's box of stencil cells in unwrapped indices (near - 1 .. near + 1 per axis)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int lo = valid ? near[a] - 1 : 0x7fffffff, hi = valid ? near[a] + 1 : -0x7fffffff;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o, 64));
            hi = max(hi, __shfl_xor(hi, o, 64));
        }
        if (lane == 0) {
            wred[a][wid] = lo;
            wred[3 + a][wid] = hi;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int v = wred[threadIdx.x][0];
        for (int w = 1; w < INTERP_THREADS / 64; w++)
            v = threadIdx.x < 3 ? min(v, wred[threadIdx.x][w]) : max(v, wred[threadIdx.x][w]);
        box[threadIdx.x] = v;
    }
    __syncthreads();
    const int bx = box[0], by = box[1], bz = box[2];
    const long long ex = (long long)box[3] - bx + 1, ey = (long long)box[4] - by + 1, ez = (long long)box[5] - bz + 1;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (ex * ey * ez <= INTERP_LDS) {  // block-uniform
        // rows (b, c) of the box to groups of 16 threads, cells of a row to the 16 lanes of the group:
        // one division per row, not per cell, and the periodic wrap by compares
        const int nrow = (int)(ey * ez), exi = (int)ex, eyi = (int)ey;
        for (int r = threadIdx.x >> 4; r < nrow; r += INTERP_THREADS >> 4) {
            const int b = r % eyi, c = r / eyi;
            const size_t rowg = (size_t)n * ((size_t)wrapc(by + b, n) + (size_t)n * wrapc(bz + c, n));
            for (int a = threadIdx.x & 15; a < exi; a += 16) sg[a + exi * r] = grad[rowg + wrapc(bx + a, n)];
        }
        __syncthreads();
        if (!valid) return;
        const int xi[3] = {near[0] - bx, near[0] + sgn[0] - bx, near[0] - sgn[0] - bx};
        const int yi[3] = {near[1] - by, near[1] + sgn[1] - by, near[1] - sgn[1] - by};
        const int zi[3] = {near[2] - bz, near[2] + sgn[2] - bz, near[2] - sgn[2] - bz};
        interp_sum(wx, wy, wz, xi, yi, zi, (size_t)ex, (size_t)(ex * ey), sg, fx, fy, fz);
    } else {
        if (!valid) return;
        const int xi[3] = {wrapc(near[0], n), wrapc(near[0] + sgn[0], n), wrapc(near[0] - sgn[0], n)};
        const int yi[3] = {wrapc(near[1], n), wrapc(near[1] + sgn[1], n), wrapc(near[1] - sgn[1], n)};
        const int zi[3] = {wrapc(near[2], n), wrapc(near[2] + sgn[2], n), wrapc(near[2] - sgn[2], n)};
        interp_sum(wx, wy, wz, xi, yi, zi, (size_t)n, (size_t)n * n, grad, fx, fy, fz);
    }
    vx[p] += (double)(fact2 * fx);
    vy[p] += (double)(fact2 * fy);
    vz[p] += (double)(fact2 * fz);
}

__global__ void interp_kick_f64(const double* __restrict__ px, const double* __restrict__ py,
                                const double* __restrict__ pz, double* __restrict__ vx, double* __restrict__ vy,
                                double* __restrict__ vz, size_t np, int n, const double4* __restrict__ grad,
                                const double* __restrict__ fac) {
    const double fact2 = fac[1];
    size_t p = (size_t)lagpm_wave::xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int ix, iy, iz, sx, sy, sz;
    double wx[3], wy[3], wz[3];
    tsc_axis_d(px[p], ix, sx, wx);
    tsc_axis_d(py[p], iy, sy, wy);
    tsc_axis_d(pz[p], iz, sz, wz);
    int xi[3] = {wrapc(ix, n), wrapc(ix + sx, n), wrapc(ix - sx, n)};
    int yi[3] = {wrapc(iy, n), wrapc(iy + sy, n), wrapc(iy - sy, n)};
    int zi[3] = {wrapc(iz, n), wrapc(iz + sz, n), wrapc(iz - sz, n)};
    double fx = 0, fy = 0, fz = 0;
    for (int c = 0; c < 3; c++)
        for (int b = 0; b < 3; b++) {
            size_t row = (size_t)n * ((size_t)yi[b] + (size_t)n * zi[c]);
            for (int a = 0; a < 3; a++) {
                double wt = wx[a] * wy[b] * wz[c];
                double4 g = grad[row + xi[a]];
                fx += wt * g.x; fy += wt * g.y; fz += wt * g.z;
            }
        }
    vx[p] += fact2 * fx;
    vy[p] += fact2 * fy;
    vz[p] += fact2 * fz;
}

unsigned grid_for(size_t work, int threads, unsigned cap = 65535u * 4) {
    size_t g = (work + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// ---- one-time check of each new FFT plan pair on the device ------------------------------
// mode 0: a single Fourier mode (1, 2, 3); mode 1: a hashed pseudo-random field in [-1, 1)
template <typename T>
__global__ void fftchk_fill(T* __restrict__ xp, int n, int mode) {
    const size_t m = (size_t)n * n * n, mx = 2 * (size_t)(n / 2 + 1);
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = q % n, j = (q / n) % n, k = q / ((size_t)n * n);
        const size_t at = mx * (q / n) + i;  // padded position of element q
        if (mode == 0) {
            const size_t ph = (i + 2 * j + 3 * k) % (size_t)n;
            xp[at] = (T)cos(2.0 * M_PI * (double)ph / (double)n);
        } else {
            uint64_t h = (uint64_t)q * 0x9E3779B97F4A7C15ull;
            h ^= h >> 29; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 32;
            xp[at] = (T)((double)(h >> 11) * (1.0 / 4503599627370496.0) - 1.0);
        }
    }
}
// wave maximum of a non-negative value, then one atomic per wave on its bit pattern (the
// order of non-negative doubles is the order of their bits)
__device__ inline void atomic_max_nonneg(double* a, double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long*)a, (unsigned long long)__double_as_longlong(v));
}
// max |spec - expected| / (n^3 / 2) for the single mode of fftchk_fill(mode 0)
template <typename T>
__global__ void fftchk_mode(const T* __restrict__ spec, int n, double* __restrict__ err) {
    const size_t h = n / 2 + 1, K = h * n * n;
    const double half = 0.5 * (double)n * n * n;
    double e = 0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < K; q += (size_t)gridDim.x * blockDim.x) {
        const size_t i = q % h, j = (q / h) % n, k = q / (h * n);
        const double want = (i == 1 && j == 2 % (size_t)n && k == 3 % (size_t)n) ? half : 0.0;
        e = fmax(e, fmax(fabs((double)spec[2 * q] - want), fabs((double)spec[2 * q + 1])) / half);
    }
    atomic_max_nonneg(err, e);
}
// max |a * sa - b| over cnt elements; row > 0: only the first `row` of every `pitch` elements
template <typename T>
__global__ void fftchk_cmp(const T* __restrict__ a, const T* __restrict__ b, size_t cnt, double sa, size_t row, size_t pitch,
                           double* __restrict__ err) {
    double e = 0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += (size_t)gridDim.x * blockDim.x) {
        if (row && q % pitch >= row) continue;
        e = fmax(e, fabs((double)a[q] * sa - (double)b[q]));
    }
    atomic_max_nonneg(err, e);
}

int fft_status(lagpm_ctx* ctx, rocfft_status s, const char* what) {
    if (s == rocfft_status_success) return LAGPM_OK;
    return lagpm_fail(ctx, LAGPM_E_FFT, std::string("rocFFT ") + what + " failed, status " + std::to_string((int)s));
}

int make_exec(lagpm_ctx* ctx, fft_plans* fp, hipStream_t st, rocfft_execution_info* info, void** work);
int check_plans(lagpm_ctx* ctx, int n, int prec, fft_plans* fp, hipStream_t st);

int get_plans(lagpm_ctx* ctx, int n, int prec, hipStream_t st, fft_plans** out) {
    auto key = std::make_pair(n, prec);
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        *out = &it->second;
        return LAGPM_OK;
    }
    fft_plans fp;
    size_t lengths[3] = {(size_t)n, (size_t)n, (size_t)n};
    rocfft_precision rp = prec == LAGPM_PREC_FP64 ? rocfft_precision_double : rocfft_precision_single;
    // In place on the padded mesh. (The out-of-place 3-D real transforms of this rocFFT gave
    // wrong, run-to-run different results for n = 108 on MI355X; in place they are exact.)
    const size_t h = (size_t)n / 2 + 1;
    size_t rs[3] = {1, 2 * h, 2 * h * (size_t)n}, cs[3] = {1, h, h * (size_t)n};
    const size_t rdist = 2 * h * (size_t)n * n, cdist = h * (size_t)n * n;
    rocfft_plan_description dfw = nullptr, dinv = nullptr;
    RCHK(fft_status(ctx, rocfft_plan_description_create(&dfw), "description_create"));
    RCHK(fft_status(ctx, rocfft_plan_description_create(&dinv), "description_create"));
    int rcd = fft_status(ctx, rocfft_plan_description_set_data_layout(dfw, rocfft_array_type_real,
                                                                      rocfft_array_type_hermitian_interleaved, nullptr,
                                                                      nullptr, 3, rs, rdist, 3, cs, cdist), "data_layout");
    if (rcd == LAGPM_OK)
        rcd = fft_status(ctx, rocfft_plan_description_set_data_layout(dinv, rocfft_array_type_hermitian_interleaved,
                                                                      rocfft_array_type_real, nullptr, nullptr, 3, cs,
                                                                      cdist, 3, rs, rdist), "data_layout");
    if (rcd == LAGPM_OK)
        rcd = fft_status(ctx, rocfft_plan_create(&fp.fwd, rocfft_placement_inplace, rocfft_transform_type_real_forward, rp,
                                                 3, lengths, 1, dfw), "plan_create(r2c)");
    if (rcd == LAGPM_OK)
        rcd = fft_status(ctx, rocfft_plan_create(&fp.inv, rocfft_placement_inplace, rocfft_transform_type_real_inverse, rp,
                                                 3, lengths, 1, dinv), "plan_create(c2r)");
    rocfft_plan_description_destroy(dfw);
    rocfft_plan_description_destroy(dinv);
    RCHK(rcd);
    size_t w1 = 0, w2 = 0;
    RCHK(fft_status(ctx, rocfft_plan_get_work_buffer_size(fp.fwd, &w1), "work_size"));
    RCHK(fft_status(ctx, rocfft_plan_get_work_buffer_size(fp.inv, &w2), "work_size"));
    fp.work_bytes = w1 > w2 ? w1 : w2;
    auto res = ctx->plans.emplace(key, fp);
    fft_plans* mine = &res.first->second;
    const int rc = check_plans(ctx, n, prec, mine, st);
    if (rc != LAGPM_OK) {
        rocfft_plan_destroy(mine->fwd);
        rocfft_plan_destroy(mine->inv);
        ctx->plans.erase(key);
        return rc;
    }
    *out = mine;
    return LAGPM_OK;
}

// The library trusts a plan pair only after it has transformed a single Fourier mode exactly,
// given the same spectrum twice for the same input, and inverted a pseudo-random field.
template <typename T>
int check_plans_t(lagpm_ctx* ctx, int n, fft_plans* fp, hipStream_t st, double tol) {
    const size_t m = (size_t)n * n * n, h = (size_t)n / 2 + 1, K = h * n * n, mx = 2 * h;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    RCHK(make_exec(ctx, fp, st, &info, &work));
    T *a = nullptr, *b = nullptr, *ref = nullptr;  // padded meshes (2K reals each)
    double* err = nullptr;
    int rc = LAGPM_OK;
    if (hipMalloc(&a, 2 * K * sizeof(T)) != hipSuccess || hipMalloc(&b, 2 * K * sizeof(T)) != hipSuccess ||
        hipMalloc(&ref, 2 * K * sizeof(T)) != hipSuccess || hipMalloc(&err, 4 * sizeof(double)) != hipSuccess)
        rc = lagpm_fail(ctx, LAGPM_E_NOMEM, "FFT check buffers");
    double e[4] = {0, 0, 0, 0};
    const unsigned g = grid_for(m, 256);
    auto run = [&](rocfft_plan plan, T* x) {
        void* ip[1] = {x};
        return fft_status(ctx, rocfft_execute(plan, ip, nullptr, info), "execute(check)");
    };
    if (rc == LAGPM_OK && hipMemsetAsync(err, 0, 4 * sizeof(double), st) != hipSuccess) rc = LAGPM_E_HIP;
    if (rc == LAGPM_OK) {
        hipLaunchKernelGGL(fftchk_fill<T>, dim3(g), dim3(256), 0, st, a, n, 0);
        rc = run(fp->fwd, a);
    }
    if (rc == LAGPM_OK) {
        hipLaunchKernelGGL(fftchk_mode<T>, dim3(g), dim3(256), 0, st, a, n, err);
        hipLaunchKernelGGL(fftchk_fill<T>, dim3(g), dim3(256), 0, st, a, n, 1);
        hipLaunchKernelGGL(fftchk_fill<T>, dim3(g), dim3(256), 0, st, b, n, 1);
        hipLaunchKernelGGL(fftchk_fill<T>, dim3(g), dim3(256), 0, st, ref, n, 1);
        rc = run(fp->fwd, a);
    }
    if (rc == LAGPM_OK) rc = run(fp->fwd, b);
    if (rc == LAGPM_OK) {
        hipLaunchKernelGGL(fftchk_cmp<T>, dim3(g), dim3(256), 0, st, a, b, 2 * K, 1.0, (size_t)0, (size_t)1, err + 1);
        rc = run(fp->inv, b);
    }
    if (rc == LAGPM_OK) {
        hipLaunchKernelGGL(fftchk_cmp<T>, dim3(g), dim3(256), 0, st, b, ref, 2 * K, 1.0 / (double)m, (size_t)n, mx, err + 2);
        rc = lagpm_d2h(ctx, e, err, sizeof(e), st);
    }
    (void)hipFree(a); (void)hipFree(b); (void)hipFree(ref); (void)hipFree(err);
    if (info) rocfft_execution_info_destroy(info);
    if (work) (void)hipFree(work);
    if (rc != LAGPM_OK) return rc;
    if (!(e[0] <= tol && e[1] == 0.0 && e[2] <= tol))
        return lagpm_fail(ctx, LAGPM_E_FFT,
                          "rocFFT plan for n=" + std::to_string(n) + " failed its check (single mode error " + std::to_string(e[0]) +
                              ", repeat difference " + std::to_string(e[1]) + ", round trip error " + std::to_string(e[2]) +
                              "); this mesh size cannot be used with this rocFFT");
    return LAGPM_OK;
}

int check_plans(lagpm_ctx* ctx, int n, int prec, fft_plans* fp, hipStream_t st) {
    return prec == LAGPM_PREC_FP64 ? check_plans_t<double>(ctx, n, fp, st, 1e-9) : check_plans_t<float>(ctx, n, fp, st, 1e-4);
}

// Execution info + work buffer for a plan pair on a stream (caller owns both).
int make_exec(lagpm_ctx* ctx, fft_plans* fp, hipStream_t st, rocfft_execution_info* info, void** work) {
    *info = nullptr;
    *work = nullptr;
    RCHK(fft_status(ctx, rocfft_execution_info_create(info), "execution_info_create"));
    if (fp->work_bytes) {
        HIPCHK(ctx, hipMalloc(work, fp->work_bytes));
        RCHK(fft_status(ctx, rocfft_execution_info_set_work_buffer(*info, *work, fp->work_bytes), "set_work_buffer"));
    }
    RCHK(fft_status(ctx, rocfft_execution_info_set_stream(*info, st), "set_stream"));
    return LAGPM_OK;
}

}  // namespace

// ---- CorrMeasure: binned power spectrum of the density contrast, Cosmos/Correlation.c:33-77 ----
// One thread per half-complex mode with i < n/2. Each block accumulates its modes' linear bin
// weights in LDS and adds them to the global bins once.
__device__ inline double tsc_sinc3(double m, int n) {
    if (m == 0.0) return 1.0;
    const double x = M_PI * m / (double)n;
    const double s = sin(x) / x;
    return s * s * s;
}
template <typename T>
__global__ void __launch_bounds__(256) pk_bin_kernel(const T* __restrict__ spec, int n, double* __restrict__ pk,
                                                     double* __restrict__ npk) {
    extern __shared__ double sh[];  // 2 * (n + 1)
    double* lpk = sh;
    double* lnpk = sh + n + 1;
    for (int b = threadIdx.x; b <= n; b += blockDim.x) lpk[b] = lnpk[b] = 0.0;
    __syncthreads();
    const int h = n / 2 + 1, ngh = n / 2;
    const size_t modes = (size_t)ngh * n * n;
    const double rngm6 = 1.0 / ((double)n * n * n * (double)n * n * n);
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < modes; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ngh);
        const size_t r = q / ngh;
        const int j = (int)(r % n), k = (int)(r / n);
        const int wj = j > ngh ? j - n : j, wk = k > ngh ? k - n : k;
        const double w = tsc_sinc3((double)i, n) * tsc_sinc3((double)wj, n) * tsc_sinc3((double)wk, n);
        const size_t e = (size_t)i + (size_t)h * ((size_t)j + (size_t)n * k);
        const double re = (double)spec[2 * e], im = (double)spec[2 * e + 1];
        const double d = (re * re + im * im) * rngm6 / (w * w);
        const double wave = sqrt((double)i * i + (double)wj * wj + (double)wk * wk);
        const int b1 = (int)wave;
        double t = wave - (double)b1;
        const int b2 = b1 + (t > 0 ? 1 : -1);
        t = fabs(t);
        if (b1 <= n) {
            atomicAdd(&lpk[b1], (1 - t) * d);
            atomicAdd(&lnpk[b1], 1 - t);
        }
        if (b2 >= 0 && b2 <= n) {
            atomicAdd(&lpk[b2], t * d);
            atomicAdd(&lnpk[b2], t);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b <= n; b += blockDim.x) {
        if (lnpk[b] != 0.0) {
            atomicAdd(&pk[b], lpk[b]);
            atomicAdd(&npk[b], lnpk[b]);
        }
    }
}

// Plans (with their one-time check) and the box's execution info exist before a step is
// captured: both synchronise or allocate, which a capture must not contain.
int lagpm_pm_prepare(lagpm_ctx* ctx, lagpm_box* box) {
    fft_plans* fp = nullptr;
    RCHK(get_plans(ctx, box->p.n, box->p.precision, box->st, &fp));
    if (!box->fft_info) RCHK(make_exec(ctx, fp, box->st, &box->fft_info, &box->fft_work));
    RCHK(lagpm_green_prepare(ctx, box->st, box->p.n, box->p.efold, box->p.deconvolve));
    return LAGPM_OK;
}

__global__ void zero_f64_kernel(double* __restrict__ a, size_t m) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0;
}

// Bins of the allocation-free power spectrum (lagpm_power_spectrum_noalloc): pk and npk, n + 1
// each, for n up to the library's largest mesh (4096): 65,552 bytes on every device the library
// runs on, held until the process ends. In host code `scratch` is only a stand-in name: the
// device address of each GPU's copy comes from hipGetSymbolAddress (lagpm_pk_scratch_init).
static __device__ __align__(256) double scratch[LAGPM_PK_SCRATCH_DOUBLES];

namespace {
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of that device's scratch
}  // namespace

int lagpm_pk_scratch_init(lagpm_ctx* ctx) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(ctx->device);
    if (it == g_scratch_ptr.end()) {
        // the device is current (lagpm_create set it), so this is that device's copy
        double* ptr = nullptr;
        HIPCHK(ctx, hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return lagpm_fail(ctx, LAGPM_E_HIP, "no device address for the power-spectrum scratch");
        it = g_scratch_ptr.emplace(ctx->device, ptr).first;
    }
    ctx->pk_scratch = it->second;
    return LAGPM_OK;
}

namespace {
// r2c of the resident density (rho) into box->spec with the box's execution info, then the
// binned |delta_k|^2 / n^6 into d (2(n+1) doubles: pk, then npk), copied to the host
int pm_power_into(lagpm_ctx* ctx, lagpm_box* box, fft_plans* fp, double* d, double* pk_host, double* npk_host) {
    const int n = box->p.n;
    RCHK(fft_status(ctx, rocfft_execution_info_set_stream(box->fft_info, box->st), "set_stream"));
    void* in[1] = {box->rho};
    RCHK(fft_status(ctx, rocfft_execute(fp->fwd, in, nullptr, box->fft_info), "execute(r2c)"));
    const size_t nbins = 2 * (size_t)(n + 1);
    // a fill kernel rather than a memset: every call starts from zero bins, whatever the last
    // call (of any mesh size) left there
    hipLaunchKernelGGL(zero_f64_kernel, dim3(grid_for(nbins, 256)), dim3(256), 0, box->st, d, nbins);
    HIPCHK(ctx, hipGetLastError());
    const size_t modes = (size_t)(n / 2) * n * n;
    const size_t shm = nbins * sizeof(double);
    if (box->p.precision == LAGPM_PREC_FP64)
        hipLaunchKernelGGL(pk_bin_kernel<double>, dim3(grid_for(modes, 256, 4096)), dim3(256), shm, box->st,
                           (const double*)box->spec, n, d, d + n + 1);
    else
        hipLaunchKernelGGL(pk_bin_kernel<float>, dim3(grid_for(modes, 256, 4096)), dim3(256), shm, box->st,
                           (const float*)box->spec, n, d, d + n + 1);
    HIPCHK(ctx, hipGetLastError());
    RCHK(lagpm_d2h(ctx, pk_host, d, (n + 1) * sizeof(double), box->st));
    RCHK(lagpm_d2h(ctx, npk_host, d + n + 1, (n + 1) * sizeof(double), box->st));
    return LAGPM_OK;
}
}  // namespace

int lagpm_pm_power(lagpm_ctx* ctx, lagpm_box* box, double* pk_host, double* npk_host) {
    const int n = box->p.n;
    RCHK(lagpm_pm_prepare(ctx, box));
    fft_plans* fp = nullptr;
    RCHK(get_plans(ctx, n, box->p.precision, box->st, &fp));
    double* d = nullptr;
    HIPCHK(ctx, hipMalloc(&d, 2 * (size_t)(n + 1) * sizeof(double)));
    const int rc = pm_power_into(ctx, box, fp, d, pk_host, npk_host);
    (void)hipStreamSynchronize(box->st);
    (void)hipFree(d);
    return rc;
}

// The same measurement with the bins in the device's static scratch: no allocation of any kind.
// The box's FFT plans and execution info must exist (any earlier PM kick or power spectrum of a
// box of this size and precision on this context creates them), since creating them allocates.
int lagpm_pm_power_noalloc(lagpm_ctx* ctx, lagpm_box* box, double* pk_host, double* npk_host) {
    const int n = box->p.n;
    if (2 * (size_t)(n + 1) > (size_t)LAGPM_PK_SCRATCH_DOUBLES)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "mesh of " + std::to_string(n) + " cells per side exceeds the power-spectrum "
                                              "scratch (n <= " + std::to_string(LAGPM_PK_SCRATCH_DOUBLES / 2 - 1) + ")");
    if (!ctx->pk_scratch) return lagpm_fail(ctx, LAGPM_E_STATE, "power-spectrum scratch not initialised");
    auto it = ctx->plans.find(std::make_pair(n, box->p.precision));
    if (it == ctx->plans.end() || !box->fft_info)
        return lagpm_fail(ctx, LAGPM_E_STATE, "no FFT plan or execution info for this box yet (creating them allocates): "
                                              "run one PM kick or lagpm_power_spectrum on it first");
    return pm_power_into(ctx, box, &it->second, ctx->pk_scratch, pk_host, npk_host);
}

int lagpm_box_fft_release(lagpm_box* box) {
    if (box->fft_info) rocfft_execution_info_destroy(box->fft_info);
    if (box->fft_work) (void)hipFree(box->fft_work);
    box->fft_info = nullptr;
    box->fft_work = nullptr;
    return LAGPM_OK;
}

int lagpm_fft_release(lagpm_ctx* ctx) {
    for (auto& kv : ctx->plans) {
        fft_plans& fp = kv.second;
        if (fp.fwd) rocfft_plan_destroy(fp.fwd);
        if (fp.inv) rocfft_plan_destroy(fp.inv);
    }
    ctx->plans.clear();
    for (auto& kv : ctx->green_tabs) (void)hipFree(kv.second);
    ctx->green_tabs.clear();
    return LAGPM_OK;
}

// Morton level whose blocks are at most 13 cells wide (the LDS box of a deposit chunk).
static int deposit_block_level(int n) {
    int lvl = 0;
    while (lvl < LAGPM_KEY_BITS && (double)n / (double)(1 << lvl) > 13.0) lvl++;
    return lvl;
}

size_t lagpm_deposit_blocks(int n) { return (size_t)1 << (3 * deposit_block_level(n)); }

size_t lagpm_deposit_chunk_cap(int n, size_t np) {
    return np / DEP_CHUNK + std::min(np, lagpm_deposit_blocks(n)) + 2;
}

int lagpm_pm_deposit(lagpm_ctx* ctx, lagpm_box* box) {
    const int n = box->p.n;
    const size_t m = (size_t)n * n * n;
    const bool f64 = box->p.precision == LAGPM_PREC_FP64;
    hipStream_t st = box->st;
    const double scale = std::ldexp(1.0, box->scale_bits);
    if (!box->np) {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(m, 256)), dim3(256), 0, st, box->acc, m);
        HIPCHK(ctx, hipGetLastError());
    } else {
        // chunk list from the sorted keys (the caller sorted the particles); the chunk count
        // stays on the device and the deposit grid covers its upper bound. Its scratch is its own
        // (dep_blk, dep_scan_tmp): a step runs the deposit beside the tree build, which uses
        // ncount, noff and scan_tmp.
        const int lvl = deposit_block_level(n);
        const int sh = 3 * (LAGPM_KEY_BITS - lvl);
        const uint32_t nblk = (uint32_t)lagpm_deposit_blocks(n);
        const size_t np = box->np;
        uint32_t* bstart = box->dep_blk;
        uint32_t* bcnt = bstart + (nblk + 1);
        uint32_t* boff = bcnt + (nblk + 1);
        hipLaunchKernelGGL(dep_blocks_kernel, dim3(grid_for(std::max((size_t)nblk + 1, m), 256)), dim3(256), 0, st, box->key,
                           np, sh, nblk, bstart, bcnt, box->acc, m);
        HIPCHK(ctx, hipGetLastError());
        uint32_t* nchunk = box->dsz + DSZ_CHUNKS;
        RCHK(lagpm_exclusive_scan_u32(ctx, st, box->dep_scan_tmp, bcnt, boff, nblk, nchunk));
        hipLaunchKernelGGL(dep_chunks_kernel, dim3(grid_for((size_t)nblk, 256)), dim3(256), 0, st, bstart, bcnt, boff, nblk,
                           np, box->dep_start, nchunk);
        HIPCHK(ctx, hipGetLastError());
        // persistent blocks, a multiple of the 8 XCDs, as many as are resident at once
        const unsigned cap = (unsigned)(box->dep_cap - 1);
        const size_t resident = (size_t)(std::max(ctx->num_cu, 8) * DEP_BLOCKS_PER_CU + 7) / 8 * 8;
        const unsigned grid = (unsigned)std::min(resident, ((size_t)cap + 7) / 8 * 8);
        // contributions are at most mass (a product of TSC weights is below 1): the 2^52 rounding
        // applies when mass 2^s stays below 2^51
        const bool magic = std::ldexp((double)box->p.mass, box->scale_bits) < 0x1p51;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(grid), dim3(DEP_THREADS), 0, st, box->x, box->y, box->z, box->dep_start, nchunk,
                               cap, np, n, (float)box->p.mass, scale, box->acc);
        };
        if (f64) {
            if (magic) launch(deposit_kernel<true, true>);
            else launch(deposit_kernel<true, false>);
        } else {
            if (magic) launch(deposit_kernel<false, true>);
            else launch(deposit_kernel<false, false>);
        }
        HIPCHK(ctx, hipGetLastError());
    }
    const double inv_scale = std::ldexp(1.0, -box->scale_bits);
    const double mean = 1.0;  // the FFT grid starts at -1 before the fold (Cosmos/pmmain.c:304)
    if (f64)
        hipLaunchKernelGGL(convert_kernel<double>, dim3(grid_for(m, 256)), dim3(256), 0, st, box->acc, n, inv_scale,
                           mean, (double*)box->rho);
    else
        hipLaunchKernelGGL(convert_kernel<float>, dim3(grid_for(m, 256)), dim3(256), 0, st, box->acc, n, inv_scale,
                           mean, (float*)box->rho);
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}

// Largest Green factor table a context keeps (n <= 1024: 513^3 doubles, 1.08 GB); larger meshes
// compute the factor per cell.
constexpr size_t GREEN_TAB_MAX_BYTES = (size_t)1100 << 20;

static double* green_table_find(lagpm_ctx* ctx, int n, double efold, int deconvolve) {
    auto it = ctx->green_tabs.find(std::make_tuple(n, efold, deconvolve));
    return it == ctx->green_tabs.end() ? nullptr : it->second;
}

// Makes the context's Green factor table for (n, efold, deconvolve) if the mesh is small enough
// for one; allocates, so it runs before any captured step (lagpm_pm_prepare) or standalone call.
int lagpm_green_prepare(lagpm_ctx* ctx, hipStream_t st, int n, double efold, int deconvolve) {
    const size_t h = (size_t)n / 2 + 1, bytes = h * h * h * sizeof(double);
    if (bytes > GREEN_TAB_MAX_BYTES || green_table_find(ctx, n, efold, deconvolve)) return LAGPM_OK;
    double* tab = nullptr;
    RCHK(lagpm_dalloc_bytes(ctx, (void**)&tab, bytes));
    const double rn = n;
    hipLaunchKernelGGL(green_table_kernel, dim3(grid_for(h * h * h, 256)), dim3(256), 0, st, tab, n,
                       -std::pow(M_PI * efold / rn, 2), deconvolve);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipFree(tab);
        return lagpm_fail(ctx, LAGPM_E_HIP, std::string("green table: ") + hipGetErrorString(e));
    }
    ctx->green_tabs.emplace(std::make_tuple(n, efold, deconvolve), tab);
    return LAGPM_OK;
}

int lagpm_green_device(lagpm_ctx* ctx, hipStream_t st, int n, int prec, double efold, int deconvolve, void* spec) {
    const double rn = n;
    const double sfact = -std::pow(M_PI * efold / rn, 2);
    const double irngc = 1.0 / (rn * rn * rn);
    // workgroups of whole wavefronts covering a half-complex row of n / 2 + 1 cells
    const int h = n / 2 + 1;
    const int bs = h >= 256 ? 256 : (h + 63) / 64 * 64;
    const int segs = (h + bs - 1) / bs;
    const dim3 grid((unsigned)((size_t)n * n * segs));
    if (const double* tab = green_table_find(ctx, n, efold, deconvolve)) {
        if (prec == LAGPM_PREC_FP64)
            hipLaunchKernelGGL(green_tab_kernel<double>, grid, dim3(bs), 0, st, (double*)spec, tab, n, segs, irngc);
        else
            hipLaunchKernelGGL(green_tab_kernel<float>, grid, dim3(bs), 0, st, (float*)spec, tab, n, segs, irngc);
    } else if (prec == LAGPM_PREC_FP64) {
        hipLaunchKernelGGL(green_kernel<double>, grid, dim3(bs), 0, st, (double*)spec, n, segs, sfact, irngc, deconvolve);
    } else {
        hipLaunchKernelGGL(green_kernel<float>, grid, dim3(bs), 0, st, (float*)spec, n, segs, sfact, irngc, deconvolve);
    }
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}

int lagpm_pm_solve(lagpm_ctx* ctx, lagpm_box* box, hipEvent_t* ev) {
    fft_plans* fp = nullptr;
    RCHK(get_plans(ctx, box->p.n, box->p.precision, box->st, &fp));
    if (!box->fft_info) RCHK(make_exec(ctx, fp, box->st, &box->fft_info, &box->fft_work));
    // the box's execution info follows the stream its work is queued on
    RCHK(fft_status(ctx, rocfft_execution_info_set_stream(box->fft_info, box->st), "set_stream"));
    rocfft_execution_info info = box->fft_info;
    void* in[1] = {box->rho};  // rho, spec and phi are one padded mesh
    RCHK(fft_status(ctx, rocfft_execute(fp->fwd, in, nullptr, info), "execute(r2c)"));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[0], box->st));
    // (the table exists already when this runs inside a captured step: lagpm_pm_prepare made it)
    RCHK(lagpm_green_prepare(ctx, box->st, box->p.n, box->p.efold, box->p.deconvolve));
    RCHK(lagpm_green_device(ctx, box->st, box->p.n, box->p.precision, box->p.efold, box->p.deconvolve, box->spec));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[1], box->st));
    void* in2[1] = {box->spec};
    RCHK(fft_status(ctx, rocfft_execute(fp->inv, in2, nullptr, info), "execute(c2r)"));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[2], box->st));
    return LAGPM_OK;
}

int lagpm_pm_gradient(lagpm_ctx* ctx, lagpm_box* box) {
    const int n = box->p.n;
    // workgroups of whole wavefronts covering a row (256 threads, or the row rounded up to 64)
    const int bs = n >= 256 ? 256 : (n + 63) / 64 * 64;
    const int segs = (n + bs - 1) / bs;
    const dim3 grid((unsigned)((size_t)n * n * segs));
    if (box->p.precision == LAGPM_PREC_FP64)
        hipLaunchKernelGGL((gradient_kernel<double, double4>), grid, dim3(bs), 0, box->st, (const double*)box->phi, n, segs,
                           (double4*)box->grad);
    else
        hipLaunchKernelGGL((gradient_kernel<float, float4>), grid, dim3(bs), 0, box->st, (const float*)box->phi, n, segs,
                           (float4*)box->grad);
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}

int lagpm_pm_interp(lagpm_ctx* ctx, lagpm_box* box) {
    if (box->np == 0) return LAGPM_OK;
    const unsigned nblk = (unsigned)((box->np + 255) / 256);
    if (box->p.precision == LAGPM_PREC_FP64)
        hipLaunchKernelGGL(interp_kick_f64, dim3(nblk), dim3(256), 0, box->st, box->x, box->y, box->z, box->vx, box->vy,
                           box->vz, box->np, box->p.n, (const double4*)box->grad, (const double*)box->dfac);
    else
        hipLaunchKernelGGL(interp_kick_f32, dim3(nblk), dim3(INTERP_THREADS), 0, box->st, box->x, box->y, box->z, box->vx, box->vy,
                           box->vz, box->np, box->p.n, (const float4*)box->grad, (const double*)box->dfac);
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}
