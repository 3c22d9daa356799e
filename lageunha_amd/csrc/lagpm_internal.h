// lagpm_internal.h — device state and kernel entry points shared by the lagpm sources.
#pragma once
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lagpm.h"

#define LAGPM_NACCEPT 10     // nodes with fewer particles are never accepted (ost.h:4 MIN_CELL_PARTICLE_NUM)
#define LAGPM_KEY_BITS 16    // Morton bits per axis; octree depth <= 16
#define LAGPM_MAX_INDEX (1u << 27)  // particles per box: walk lists pack (index << 5 | neighbour) in 32 bits
#define LAGPM_MAX_NODES (1u << 26)  // nodes per box: walk stacks pack (node << 5 | neighbour) in a non-negative int

// Doubles in the static power-spectrum scratch of each device: pk and npk bins for n <= 4096.
#define LAGPM_PK_SCRATCH_DOUBLES (2 * (4096 + 1))

struct fft_plans {
    rocfft_plan fwd = nullptr, inv = nullptr;
    size_t work_bytes = 0;
};

// Force-law tables normalised to n = 1 (the law scales as 1/n^2), shared by every box of a
// context that has the same law, softening and split scale.
// Two bin layouts: the reference's 4096 bins uniform in r over [0, 6) (LAW_PMFIT, so reference
// mode interpolates exactly as OST/Model4TreeForce.c:34-46), and for LAW_ERFC bins uniform in
// log(lam r^2), 256 per octave over lam r^2 in [2^-14, 64) with lam = 32 / rc^2, indexed by the
// bits of the FP32 lam r^2 (no square root in the pair loop; about ten times smaller
// interpolation error; the cutoff r^2 = rc^2 is the bin edge 32, beyond which bins are zero).
#define LAGPM_LOG_SHIFT 15                 // 23 - 8 mantissa bits kept: 256 bins per octave
#define LAGPM_LOG_BASE 28928u              // bits(2^-14) >> 15
#define LAGPM_LOG_BINS 5120                // 20 octaves: r^2 < 64
struct law_tables {
    int logr2 = 0;            // 1: log r^2 bins, 0: uniform r bins
    int nbins = 0;
    double rc2_eff = 0;       // squared cutoff the table realises
    float lam = 1.0f;         // log bins: r^2 is binned as lam * r^2, lam = 32 / rc^2 (the cutoff is a bin edge)
    float2* tab0 = nullptr;   // (d0, s0) per bin: particle term
    float4* tab12 = nullptr;  // (d1, s1, d2, s2) per bin: quadrupole terms
    double* tabd = nullptr;   // 6 per bin: d0 s0 d1 s1 d2 s2 (FP64 walk)
};

struct walk_box;  // device view of one box for the walk (tree.hip)

// XCDs of one MI355X in SPX mode: the walk splits its groups into this many ranges.
constexpr int LAGPM_WALK_XCDS = 8;

struct lagpm_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;            // default stream of the context
    std::vector<hipStream_t> pool;           // per-box streams of batched steps
    std::string err;
    std::map<std::pair<int, int>, fft_plans> plans;  // (n, precision) -> plans
    std::map<std::tuple<int, double, double, double>, law_tables> tables;  // (law, epsilon, efold, rsphere)
    // walk resources (every walk runs on `stream`, so one set serves all boxes)
    int* walk_stack = nullptr;
    size_t walk_stack_entries = 0;
    unsigned long long* walk_ticket = nullptr;  // LAGPM_WALK_XCDS group tickets, one per XCD range
    walk_box* walk_boxes = nullptr;          // descriptor array of the last launch
    uint32_t* walk_prefix = nullptr;         // group prefix over the descriptors
    size_t walk_box_cap = 0;
    bool use_graphs = true;                  // capture box step fronts in hipGraphs
    bool walk_census = false;                // FP32 walks count the pairs within the cutoff
    int batch_stages = 5;                    // groups of a pipelined batch step (1: no pipeline)
    int stage_reserve_cu = 32;               // CUs a non-final group's walk leaves to the next fronts
    hipEvent_t ev[4] = {};
    double* pk_scratch = nullptr;  // this device's static power-spectrum scratch (device address)
    // Green factor tables, (n, efold, deconvolve) -> (n/2+1)^3 doubles (pm.hip)
    std::map<std::tuple<int, double, int>, double*> green_tabs;
    // pinned host buffers for the end-of-step readbacks: one record per box of a batch, and one
    // for a lone box's finish (so a redo inside a batch does not overwrite the batch's records)
    void* readback = nullptr;
    size_t readback_count = 0;
    void* readback1 = nullptr;
};

// Tunables of the tree correction (chosen per box; the autotuner may override).
struct tree_tune {
    int nleaf = 16;        // a node with more particles than this is split
    int group = 64;        // maximal target group (one wavefront)
    int runs = 0;          // 0: tree groups; 1: runs of `group` particles of a chaining cell; 2: packed tree groups
    int walk_waves = 0;    // persistent walk waves (0: auto)
};

// Device-side sizes of a box's tree and deposit, written by the kernels that produce them so
// that a step never waits for the host. Indices into lagpm_box::dsz.
// device size words: node/group/deposit-chunk counts, the overflow flag, and for the tree build
// the nodes per level (DSZ_LVL + level) and their offsets in the level order (DSZ_LOFF + level)
enum {
    DSZ_NODES = 0,
    DSZ_GROUPS = 1,
    DSZ_CHUNKS = 2,
    DSZ_OVERFLOW = 3,
    DSZ_SCRATCH = 4,  // scan totals nobody reads
    DSZ_DEEPER = 5,   // the build met a node to split at its depth limit (the tree is redone)
    DSZ_WALKOVF = 6,  // a walk stack overflowed: that step's drift is skipped (the error is returned)
    DSZ_LVL = 8,
    DSZ_LOFF = DSZ_LVL + LAGPM_KEY_BITS + 1,
    DSZ_LCUR = DSZ_LOFF + LAGPM_KEY_BITS + 1,
    DSZ_COUNT = DSZ_LCUR + LAGPM_KEY_BITS + 1
};

// A box's tree sizes and interaction counters, read back at the end of a step.
struct tree_readback {
    uint32_t d[DSZ_COUNT];
    unsigned long long c[4];
};

struct lagpm_box {
    lagpm_params p{};
    size_t np = 0;
    size_t bytes = 0;
    bool uploaded = false;
    bool sorted = false;
    bool tree_built = false;  // a tree build has been queued (node arrays hold a tree)
    tree_tune tune;
    // deepest level the build splits to: the full key depth until one build has run, then two
    // levels below the deepest level that build used; a build that meets a node still to split
    // at this limit flags it, and the tree is redone at full depth (levels below the tree's
    // depth cost one empty launch each per kernel)
    int depth = LAGPM_KEY_BITS;
    hipStream_t st = nullptr;  // stream this box's work is queued on (context stream or a pool stream)
    hipStream_t st2 = nullptr;  // second stream of a step front (the tree build beside the PM kick)

    // particle state, kept in the current (sorted after a sort) order. A sort gathers into the
    // other set and swaps the two, so the pointers alternate with period two (a captured step
    // graph exists per parity).
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    uint32_t* id = nullptr;  // caller index of each particle
    double* alt[6] = {};     // the other set of x, y, z, vx, vy, vz
    uint32_t* id_alt = nullptr;
    int parity = 0;
    // sort scratch
    uint64_t *key = nullptr, *key2 = nullptr;
    uint32_t *perm = nullptr, *perm2 = nullptr;
    // kick factors read by the kernels: [0] drag fact1, [1] PM kick factor, [2] drift pfact
    double* dfac = nullptr;
    // FFT execution info and work buffer owned by the box (a captured graph refers to them)
    rocfft_execution_info fft_info = nullptr;
    void* fft_work = nullptr;
    // captured front half of a step (drag .. tree build), one per particle-set parity
    hipGraphExec_t gfront[2] = {nullptr, nullptr};  // captured step front per particle-set parity
    hipEvent_t evfork = nullptr;  // fork of the step front: the tree build branches off after the sort
    hipEvent_t evsort = nullptr;  // positions gathered: the velocity gather branches off
    hipEvent_t evvel = nullptr;   // velocities and ids gathered
    hipStream_t st3 = nullptr;    // the velocity gather's stream in a lone box's step front (else st)
    bool graph_failed = false;
    uint32_t* scan_tmp = nullptr;  // scan / radix-sort scratch of this box (sort and tree build)
    size_t scan_tmp_bytes = 0;
    uint32_t* dep_scan_tmp = nullptr;  // the deposit's own scan scratch: it may run beside the tree build

    // PM meshes
    int scale_bits = 32;
    unsigned long long* acc = nullptr;  // int64 fixed-point deposit (two's complement)
    // one in-place FFT mesh, rows of n reals padded to 2(n/2+1): the density (rho), then its
    // spectrum (spec, n*n*(n/2+1) complex), then the potential (phi); three names, one buffer
    void* rho = nullptr;
    void* spec = nullptr;
    void* phi = nullptr;
    void* grad = nullptr;                // n^3 x 4 real (fx, fy, fz, phi)
    uint32_t* dep_start = nullptr;       // deposit chunk starts (+ sentinel)
    uint32_t* dep_blk = nullptr;         // per Morton block: first particle, chunk count, chunk offset
    size_t dep_cap = 0;

    // tree
    int level_cell = 0;       // chaining level L: 2^L cells per axis, width n/2^L >= rsphere
    int mcell = 0;            // 2^L
    float4* pp = nullptr;     // sorted FP32 positions + mass
    uint32_t* ncount = nullptr;   // per-particle node count, then offsets
    uint32_t* noff = nullptr;
    size_t node_cap = 0;          // capacity of the node arrays (device count may not exceed it)
    int4* nd_info = nullptr;      // begin, count, nchild, level
    int* nd_parent = nullptr;
    int* nd_child = nullptr;      // 8 per node
    float4* nd_com = nullptr;     // com.xyz, mass
    float4* nd_q0 = nullptr;      // Qxx Qyy Qzz Qxy
    float4* nd_q1 = nullptr;      // Qxz Qyz trQ rmax
    double* nd_d = nullptr;       // FP64 moments: 12 per node (com3, mass, Q6, trQ, rmax)
    int* cellroot = nullptr;      // mcell^3
    // The FP32 walk's copy of the tree by record (the build's order: level by level, the children of
    // a node consecutive): links (first particle, count, first child record, children), the opening
    // test's geometry (com, rmax), the quadrupole (q0; q1 with the mass in place of rmax), the split's
    // first child record of each record, and the cell roots as records (mcell^3).
    int4* w_link = nullptr;
    float4* w_geo = nullptr;
    float4* w_q0 = nullptr;
    float4* w_q1 = nullptr;
    int* w_rfc = nullptr;
    int* w_cellroot = nullptr;
    uint32_t* gflag = nullptr;    // per-node group count, then offsets
    uint32_t* goff = nullptr;
    size_t group_cap = 0;
    int2* groups = nullptr;       // begin, count
    double* gbox = nullptr;       // FP64 group bbox: 6 per group
    int* pgroup = nullptr;        // group of each particle (FP64 plain walk)
    uint32_t* dsz = nullptr;      // DSZ_* device sizes
    unsigned long long* counters = nullptr;  // [0]=n_pp [1]=n_pc [2]=walk stack overflow
    const law_tables* tab = nullptr;
    lagpm_tree_stats stats{};
    // phase events: [0..7] pmmain (start, sort, tsc, r2c, green, c2r, grad, interp),
    //               [8..11] treecorrection (start, sort, build, walk)
    hipEvent_t ev[15] = {};  // phase events; [12] deposit start, [13] step front joined, [14] walk start
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

// Device-to-host copy of `bytes` on stream s that has completed when the call returns, whatever
// the outcome: the stream is synchronised even when queuing the copy fails, so no copy can still
// be writing into host memory the caller is about to free or leave (stack buffers included).
int lagpm_d2h(lagpm_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t s);

// ---- scans / sort (sort.hip) ----
// Scratch bytes a box needs for scans of n entries and for a radix sort of n keys.
size_t lagpm_scan_tmp_bytes(size_t n);
size_t lagpm_sort_tmp_bytes(size_t n);
// Exclusive scan of n u32 values on stream s (in and out must not alias); the total is written
// to total_dev (device memory) when not null. tmp: lagpm_scan_tmp_bytes(n) bytes.
int lagpm_exclusive_scan_u32(lagpm_ctx* ctx, hipStream_t s, uint32_t* tmp, const uint32_t* in, uint32_t* out, size_t n,
                             uint32_t* total_dev);
int lagpm_radix_sort_pairs(lagpm_ctx* ctx, hipStream_t s, uint32_t* tmp, uint64_t* key, uint64_t* key_alt, uint32_t* val,
                           uint32_t* val_alt, size_t n, int bits, bool* result_in_alt,
                           int lo_bit = 0);

// ---- PM (pm.hip) ----
int lagpm_pm_deposit(lagpm_ctx* ctx, lagpm_box* box);          // fills acc and rho
int lagpm_pm_solve(lagpm_ctx* ctx, lagpm_box* box, hipEvent_t* ev = nullptr);  // rho -> phi (r2c, Green, c2r)
int lagpm_pm_gradient(lagpm_ctx* ctx, lagpm_box* box);         // phi -> grad
int lagpm_pm_interp(lagpm_ctx* ctx, lagpm_box* box);  // v += dfac[1] * F(grad)
int lagpm_green_device(lagpm_ctx* ctx, hipStream_t s, int n, int prec, double efold, int deconvolve, void* spec);
// the context's Green factor table for (n, efold, deconvolve); allocates (call before capture)
int lagpm_green_prepare(lagpm_ctx* ctx, hipStream_t s, int n, double efold, int deconvolve);
size_t lagpm_deposit_chunk_cap(int n, size_t np);
size_t lagpm_deposit_blocks(int n);  // Morton blocks of the deposit's chunking
int lagpm_fft_release(lagpm_ctx* ctx);
int lagpm_box_fft_release(lagpm_box* box);
int lagpm_pm_prepare(lagpm_ctx* ctx, lagpm_box* box);
int lagpm_pm_power(lagpm_ctx* ctx, lagpm_box* box, double* pk_host, double* npk_host);
// the same with the bins in the device's static scratch; allocates nothing (LAGPM_E_STATE when the
// box's FFT plans or execution info do not exist yet)
int lagpm_pm_power_noalloc(lagpm_ctx* ctx, lagpm_box* box, double* pk_host, double* npk_host);
// looks up (once per device, with the device current) the address of the static scratch
int lagpm_pk_scratch_init(lagpm_ctx* ctx);
void lagpm_box_graphs_release(lagpm_box* box);

// ---- tree (tree.hip) ----
// With vst (a stream other than box->st): the velocities and ids are gathered on vst, and
// box->evvel marks their completion (the caller makes its stream wait for it).
int lagpm_sort_particles(lagpm_ctx* ctx, lagpm_box* box, const double* drag = nullptr, hipStream_t vst = nullptr);
// Queues the octree build on box->st; sizes stay on the device (dsz).
int lagpm_tree_build(lagpm_ctx* ctx, lagpm_box* box);
// One walk launch on ctx->stream over all groups of the given boxes (FP32 boxes sharing one
// law table); facts[i] is the kick factor of box i. Waits are the caller's business.
// reserve_cu: CUs left without a walk workgroup (for kernels of other streams running beside it)
// A sampled walk: listed particles only, each as a group of one; groups and dsz are the sample's
// own (device), out its accelerations (FP64 walk: written by the walk; FP32: gathered after it).
struct walk_sample {
    const int2* groups;
    uint32_t* dsz;
    double* out;
    size_t count;
};
int lagpm_tree_walk_f32(lagpm_ctx* ctx, lagpm_box* const* boxes, const double* facts, int nbox, int reserve_cu = 0,
                        const walk_sample* sample = nullptr);
int lagpm_tree_walk_f64(lagpm_ctx* ctx, lagpm_box* box, double fact, const walk_sample* sample = nullptr);
// the tree acceleration of the particles with caller indices ids[0..count) (device arrays; slot:
// np ints, groups: count, sdsz: DSZ_COUNT words) into acc (device, 3 per particle)
int lagpm_tree_walk_sample(lagpm_ctx* ctx, lagpm_box* box, const int64_t* ids, size_t count, int* slot, int2* groups,
                           uint32_t* sdsz, double* acc);
// Reads the device sizes and counters after the walk (synchronises box->st); grows the node
// arrays and returns LAGPM_E_OVERFLOW with *retry set when they were too small.
int lagpm_tree_finish(lagpm_ctx* ctx, lagpm_box* box, bool* retry);
// the two halves of lagpm_tree_finish: queue the copy (box->st, no wait), then, after the caller
// has waited for the stream, act on it
int lagpm_tree_readback_enqueue(lagpm_ctx* ctx, lagpm_box* box, tree_readback* dst);
int lagpm_tree_finish_with(lagpm_ctx* ctx, lagpm_box* box, const tree_readback& r, bool* retry);
int lagpm_tree_alloc(lagpm_ctx* ctx, lagpm_box* box, size_t node_cap);
double lagpm_tree_scale(const lagpm_box* box);  // FP32 tree coordinate scale (sqrt(lam) or 1)
size_t lagpm_tree_node_cap_guess(size_t np, int mcell);
int lagpm_choose_cell_level(int n, double rsphere);
size_t lagpm_walk_stack_entries(int num_cu);

// ---- host helpers (api.hip) ----
void lagpm_build_tables(const lagpm_params* p, double* diff, double* slope);
int lagpm_get_tables(lagpm_ctx* ctx, const lagpm_params* p, const law_tables** out);
int lagpm_dalloc_bytes(lagpm_ctx* ctx, void** p, size_t bytes);
