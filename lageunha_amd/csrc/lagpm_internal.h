// lagpm_internal.h — device state and kernel entry points shared by the lagpm sources.
#pragma once
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdint>
#include <map>
#include <string>

#include "../../include/lagpm.h"

#define LAGPM_NACCEPT 10     // nodes with fewer particles are never accepted (ost.h:4 MIN_CELL_PARTICLE_NUM)
#define LAGPM_KEY_BITS 16    // Morton bits per axis; octree depth <= 16
#define LAGPM_MAX_INDEX (1u << 27)  // walk lists pack (index << 5 | neighbour)

struct fft_plans {
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
};

struct lagpm_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::pair<int, int>, fft_plans> plans;  // (n, precision) -> plans
    hipEvent_t ev[8] = {};
    // device scratch for scans (grown on demand)
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
};

// Tunables of the tree correction (chosen per box; the autotuner may override).
struct tree_tune {
    int nleaf = 16;        // a node with more particles than this is split
    int group = 64;        // maximal target group (one wavefront)
    int walk_waves = 0;    // persistent walk waves (0: auto)
};

struct lagpm_box {
    lagpm_params p{};
    size_t np = 0;
    size_t bytes = 0;
    bool uploaded = false;
    bool sorted = false;
    tree_tune tune;

    // particle state, kept in the current (sorted after a sort) order
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    uint32_t* id = nullptr;  // caller index of each particle
    // sort scratch
    uint64_t *key = nullptr, *key2 = nullptr;
    uint32_t *perm = nullptr, *perm2 = nullptr;
    double* dtmp = nullptr;
    uint32_t* utmp = nullptr;

    // PM meshes
    int scale_bits = 32;
    unsigned long long* acc = nullptr;  // int64 fixed-point deposit (two's complement)
    void* rho = nullptr;                 // n^3 real (float or double)
    void* spec = nullptr;                // n*n*(n/2+1) complex
    void* phi = nullptr;                 // n^3 real
    void* grad = nullptr;                // n^3 x 4 real (fx, fy, fz, phi)
    uint32_t* dep_start = nullptr;       // deposit chunk starts (+ sentinel)
    size_t dep_cap = 0;

    // tree
    int level_cell = 0;       // chaining level L: 2^L cells per axis, width n/2^L >= rsphere
    int mcell = 0;            // 2^L
    float4* pp = nullptr;     // sorted FP32 positions + mass
    uint32_t* ncount = nullptr;   // per-particle node count, then offsets
    uint32_t* noff = nullptr;
    size_t node_cap = 0;
    size_t nnodes = 0;
    int4* nd_info = nullptr;      // begin, count, nchild, level
    int* nd_parent = nullptr;
    int* nd_child = nullptr;      // 8 per node
    float4* nd_com = nullptr;     // com.xyz, mass
    float4* nd_q0 = nullptr;      // Qxx Qyy Qzz Qxy
    float4* nd_q1 = nullptr;      // Qxz Qyz trQ rmax
    double* nd_d = nullptr;       // FP64 moments: 12 per node (com3, mass, Q6, trQ, rmax)
    int* cellroot = nullptr;      // mcell^3
    uint32_t* gflag = nullptr;    // per-node group count, then offsets
    uint32_t* goff = nullptr;
    size_t ngroups = 0;
    size_t group_cap = 0;
    int2* groups = nullptr;       // begin, count
    double* gbox = nullptr;       // FP64 group bbox: 6 per group
    int* pgroup = nullptr;        // group of each particle (FP64 plain walk)
    int* walk_stack = nullptr;
    size_t walk_stack_entries = 0;
    unsigned long long* counters = nullptr;  // [0]=n_pp [1]=n_pc [2]=group ticket [3]=overflow
    float2* tab0 = nullptr;       // (d0, s0) x 4096: particle law
    float4* tab12 = nullptr;      // (d1, s1, d2, s2) x 4096: quadrupole terms
    double* tabd = nullptr;       // 6 x 4096: d0 s0 d1 s1 d2 s2
    bool tables_ready = false;
    lagpm_tree_stats stats{};
    // phase events: [0..7] pmmain (start, sort, tsc, r2c, green, c2r, grad, interp),
    //               [8..11] treecorrection (start, sort, build, walk)
    hipEvent_t ev[12] = {};
    bool pm_timed = false, tree_timed = false;
};

// ---- error helpers ----
int lagpm_fail(lagpm_ctx* ctx, int code, const std::string& msg);
#define HIPCHK(ctx, expr)                                                                     \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return lagpm_fail((ctx), LAGPM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)
#define RCHK(expr)            \
    do {                      \
        int _r = (expr);      \
        if (_r != LAGPM_OK) return _r; \
    } while (0)

// ---- scans / sort (sort.hip) ----
int lagpm_exclusive_scan_u32(lagpm_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_host);
int lagpm_radix_sort_pairs(lagpm_ctx* ctx, uint64_t* key, uint64_t* key_alt, uint32_t* val, uint32_t* val_alt,
                           size_t n, int bits, bool* result_in_alt);

// ---- PM (pm.hip) ----
int lagpm_pm_deposit(lagpm_ctx* ctx, lagpm_box* box);          // fills acc and rho
int lagpm_pm_solve(lagpm_ctx* ctx, lagpm_box* box, hipEvent_t* ev = nullptr);  // rho -> phi (r2c, Green, c2r)
int lagpm_pm_gradient(lagpm_ctx* ctx, lagpm_box* box);         // phi -> grad
int lagpm_pm_interp(lagpm_ctx* ctx, lagpm_box* box, double fact);  // v += fact * F(grad)
int lagpm_green_device(lagpm_ctx* ctx, int n, int prec, double efold, int deconvolve, void* spec);
int lagpm_fft_release(lagpm_ctx* ctx);

// ---- tree (tree.hip) ----
int lagpm_sort_particles(lagpm_ctx* ctx, lagpm_box* box);
int lagpm_tree_build(lagpm_ctx* ctx, lagpm_box* box);
int lagpm_tree_walk(lagpm_ctx* ctx, lagpm_box* box, double fact);
int lagpm_tables_upload(lagpm_ctx* ctx, lagpm_box* box);
int lagpm_choose_cell_level(int n, double rsphere);

// ---- host helpers (api.cpp) ----
void lagpm_build_tables(const lagpm_params* p, double* diff, double* slope);
size_t lagpm_tree_node_cap(size_t np, int nleaf);
