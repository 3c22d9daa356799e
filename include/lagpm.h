/*
 * lagpm.h — C ABI of the MI355X-native TreePM gravity step.
 *
 * Drop-in boundary for the GOTPM TreePM force path of EUNHA (kjhan0606/LagEunha).
 * Every entry point is plain C: opaque handles, plain pointers and sizes, int status.
 * No HIP, torch or STL types appear here; device streams are passed as void*.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   pmmain        Cosmos/pmmain.c:390      void pmmain(SimParameters*, EvolFact)
 *   TSC           Cosmos/pmmain.c:112,188  void TSC(SimParameters*, GridInfo*)
 *   psolver       Cosmos/psolver.c:13      void psolver(SimParameters*, GridInfo*)
 *   fda4          Cosmos/fda4.c:31         void fda4(SimParameters*, GridInfo*, EvolFact)
 *   treecorrection Cosmos/treecorrection.c:225 void treecorrection(SimParameters*, DeterminedEvolFact*)
 *   i_force_spline OST/Model4TreeForce.c:48 void i_force_spline(SimParameters*)
 *   ForceLaw      ForceLaw/treeforce.c:150  void i_force_spline(int nx, float Epsilon)
 *   HubbleDragDM  Cosmos/globalT.c:102     void HubbleDragDM(SimParameters*, float)
 *   drift         Cosmos/stepevolv.c:38    void onestepforwardposition(SimParameters*)
 *   XofP decode   eunha.h:1228-1232        (GOTPM + XYZDBL 32-byte particle layout)
 *   jread/jwrite  Cosmos/CosmosRW/kjhrw.c:51,155  void jread/jwrite(SimParameters*)
 *   read_head     Params/header.c:72        void read_head(FILE*, SimParameters*)
 *   PreFoF        Cosmos/PreFoF.c:267       void PreFoF(SimParameters*)
 *
 * Error behaviour: the reference returns void and calls exit()/MPI_Finalize() on error
 * (Cosmos/treecorrection.c:240-247). Here every call returns 0 on success or a negative
 * LAGPM_E_* code, and lagpm_last_error() gives the message. Nothing here exits the process.
 */
#ifndef LAGPM_H
#define LAGPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LAGPM_ABI_VERSION 4

/* status codes */
#define LAGPM_OK          0
#define LAGPM_E_INVAL    -1   /* bad argument (sizes, pointers, parameter ranges) */
#define LAGPM_E_NOMEM    -2   /* device allocation failed / HBM budget exceeded */
#define LAGPM_E_HIP      -3   /* HIP runtime error */
#define LAGPM_E_FFT      -4   /* rocFFT error */
#define LAGPM_E_STATE    -5   /* call out of order (e.g. no particles uploaded) */
#define LAGPM_E_OVERFLOW -6   /* a bounded device structure overflowed (walk stack, node pool) */
#define LAGPM_E_IO       -7   /* a particle file could not be opened, read, parsed or written */

/* PM status of the kick, same values as PUSH/PULL/KICK of eunha.h:69-77
 * (selection in Cosmos/globalT.c:113-131 and Cosmos/fda4.c:49-63) */
#define LAGPM_PUSH 0
#define LAGPM_PULL 1
#define LAGPM_KICK 2

/* short-range force law of the tree correction */
#define LAGPM_LAW_PMFIT 0  /* reference: Plummer Newton minus fitted PM force, OST/Model4TreeForce.c:92-116 */
#define LAGPM_LAW_ERFC  1  /* Plummer Newton times the erfc complement of a Gaussian filter exp(-k^2 rs^2) */
#define LAGPM_LAW_TREEFORCE 2 /* the fitted law as ForceLaw/treeforce.c:134-186 tabulates it: softening fixed at
                                 EPSILON = 0.1 (OST/ost.h:6) whatever `epsilon` says, derivatives with a step of
                                 0.2 bin, entry 0 zero (the reference law uses 0.1 bin and entry 0 = entry 1) */

/* arithmetic of the force evaluation */
#define LAGPM_PREC_FP32 0  /* FP32 mesh/FFT/pairs, FP64 positions and velocities (fast path) */
#define LAGPM_PREC_FP64 1  /* plain FP64 kernels everywhere (baseline and accuracy reference) */

/* Kick factors; same fields and meaning as EvolFact, Cosmos/CosmosEvolFactor.h:1-6. */
typedef struct lagpm_evolfact {
    float afact, bfact, pfact;
    float fact1, fact2;
    float fact1_push, fact2_push;
    float fact1_pull, fact2_pull;
} lagpm_evolfact;

/* Per-box gravity parameters (the gravity keys of SimParameters: NX, GRV_EFOLD,
 * GRV_EPSILON, RSPHERE, THETA, eunha.h:1033-1090, plus this library's switches). */
typedef struct lagpm_params {
    int32_t n;          /* mesh cells per side; the box side is n cells (NX) */
    int32_t law;        /* LAGPM_LAW_* */
    int32_t precision;  /* LAGPM_PREC_* */
    int32_t deconvolve; /* 1: divide the Green function by the TSC window squared */
    double  mass;       /* particle mass (DM_MASS); mean density n^-3 * np * mass */
    double  efold;      /* Gaussian filter exp(-w (pi efold / n)^2) (GRV_EFOLD); r_s = efold/2 cells */
    double  epsilon;    /* Plummer softening in cells (GRV_EPSILON, default 0.1) */
    double  rsphere;    /* short-range cutoff in cells (RSPHERE) */
    double  theta;      /* opening angle (THETA); a node opens when d^2 < rmax^2/theta^2 */
    double  amax;       /* AMAX: scale factor today (1+z = amax/a) */
    double  omep;       /* Omega_matter (OMEP) */
    double  omeplam;    /* Omega_Lambda (OMEPLAM) */
} lagpm_params;

typedef struct lagpm_ctx lagpm_ctx;  /* one per device; owns streams, FFT plan cache */
typedef struct lagpm_box lagpm_box;  /* one periodic box; owns resident particle/mesh/tree state */

/* Interaction and timing counters of the last tree correction of a box. */
typedef struct lagpm_tree_stats {
    uint64_t n_pp;      /* particle-particle pairs evaluated */
    uint64_t n_pc;      /* particle-node (quadrupole) pairs evaluated */
    uint64_t n_nodes;   /* octree nodes built */
    uint64_t n_groups;  /* target groups walked */
    double   ms_build;  /* sort + build + moments, device time */
    double   ms_walk;   /* walk kernel, device time */
    int32_t  deepest;   /* deepest tree level with nodes (levels of the 16-bit-per-axis keys) */
    int32_t  depth_limit; /* deepest level the next build splits to (a deeper tree is detected
                             and rebuilt at full depth) */
    uint64_t n_pp_within; /* of n_pp, the pairs within the cutoff (FP32 walk with the census on,
                             lagpm_set_walk_census; 0 otherwise) */
} lagpm_tree_stats;

/* Device time (HIP events on the context stream) of the phases of the last pmmain and
 * treecorrection of a box; phase names follow the reference's timers (Cosmos/pmmain.c:447-522). */
typedef struct lagpm_timings {
    double ms_sort;     /* Morton keys + radix sort + permutation (0 when already sorted) */
    double ms_tsc;      /* TSC deposit + fixed-point convert (CPU(TSC)) */
    double ms_fft_r2c;  /* CPU(F-FFT) */
    double ms_green;    /* CPU(psolv) */
    double ms_fft_c2r;  /* CPU(B-FFT) */
    double ms_grad;     /* FDA4 gradient mesh */
    double ms_interp;   /* TSC interpolation + kick (with ms_grad: CPU(4FDA)) */
    double ms_tree_sort;/* sort inside treecorrection (0 when pmmain already sorted) */
    double ms_build;    /* octree build + moments + groups */
    double ms_walk;     /* walk + kick (CPU(TREE) with ms_build) */
} lagpm_timings;

/* ---- context --------------------------------------------------------------------- */
int         lagpm_create(lagpm_ctx** out, int device);
int         lagpm_destroy(lagpm_ctx* ctx);
const char* lagpm_last_error(const lagpm_ctx* ctx);
int         lagpm_abi_version(void);
/* Stream all work of this context is queued on (hipStream_t as void*). */
void*       lagpm_stream(lagpm_ctx* ctx);
int         lagpm_synchronize(lagpm_ctx* ctx);

/* ---- boxes ------------------------------------------------------------------------ */
/* Allocates device state for np particles. Fails with LAGPM_E_NOMEM beyond device memory. */
int lagpm_box_create(lagpm_ctx* ctx, const lagpm_params* p, size_t np, lagpm_box** out);
int lagpm_box_destroy(lagpm_ctx* ctx, lagpm_box* box);
int lagpm_box_params(const lagpm_box* box, lagpm_params* out);
/* Changes the box's cosmology (AMAX, OMEP, OMEPLAM), which only sets the kick factors of later
 * steps: a box and its device memory can be reused for another box of the same size and force
 * settings (a serving loop's pool). No counterpart in the reference. */
int lagpm_box_set_cosmology(lagpm_box* box, double amax, double omep, double omeplam);
size_t lagpm_box_np(const lagpm_box* box);
/* Bytes of device memory the box holds (used for HBM-budget admission). */
size_t lagpm_box_device_bytes(const lagpm_box* box);
size_t lagpm_estimate_box_bytes(const lagpm_params* p, size_t np);

/* Host SoA upload/download in the caller's particle order. Positions are wrapped into
 * [0,n). vx..vz may be NULL on upload (zero velocities). */
int lagpm_box_upload(lagpm_ctx* ctx, lagpm_box* box,
                     const double* x, const double* y, const double* z,
                     const double* vx, const double* vy, const double* vz);
int lagpm_box_download(lagpm_ctx* ctx, lagpm_box* box,
                       double* x, double* y, double* z,
                       double* vx, double* vy, double* vz);
/* Upload from the GOTPM+XYZDBL particle layout, eunha.h:372-400 (32 bytes: u64 index|flag,
 * f32 dx,dy,dz displacement from lattice site, f32 vx,vy,vz). The lattice index uses the
 * low 56 bits, x fastest: idx = i + nx*(j + nx*k); XofP = fmod(dx + i + nx, nx). */
int lagpm_box_upload_eunha(lagpm_ctx* ctx, lagpm_box* box, const void* dm_bp, size_t np, long nx);
/* Download velocities back into the same 32-byte layout (positions untouched). */
int lagpm_box_download_eunha_vel(lagpm_ctx* ctx, lagpm_box* box, void* dm_bp, size_t np);
/* Download positions and velocities into the 32-byte layout: the index/flag word of each record
 * is kept and names its lattice site; x,y,z become the displacement from that site (the periodic
 * image nearest to it, in [-n/2, n/2)), so XofP of the record gives the position back. */
int lagpm_box_download_eunha(lagpm_ctx* ctx, lagpm_box* box, void* dm_bp, size_t np);

/* ---- PreFoF halo candidates --------------------------------------------------------- */
/* PreFoF (Cosmos/PreFoF.c:267-386): friends-of-friends groups with linking length fof_link
 * (mesh units) inside each cell of width 4 alone (OST/fof.c:220-293); a group is kept unless it
 * lies inside [fof_link, 4 - fof_link) on every axis with fewer than minnum members (the
 * reference fixes FOF_MINNUM = 20, PreFoF.c:273). flags (np bytes, caller order, may be NULL)
 * gets 1 for the particles of kept groups (the reference's FoFflag); nflagged their count.
 * Uses box->np * ~93 bytes of device scratch for the call. */
int lagpm_prefof(lagpm_ctx* ctx, lagpm_box* box, float fof_link, int minnum, uint8_t* flags, size_t* nflagged);

/* ---- GOTPM particle files and records (host only, no GPU needed) --------------------- */
/* What the ASCII header of a jread/jwrite file holds (Params/params.h:49-139 keys). */
typedef struct lagpm_snapshot_info {
    long nx, ny, nz;                           /* "Nx" "Ny" "Nz" */
    long np_sum, np_dm, np_sph, np_star, np_agn; /* "Np" "DNp" "BNp" "SNp" "ANp" */
    int ptype;                                 /* "Particle Type": 0 PMTYPE (32 B), 1 TREETYPE (40 B) */
    int nspace, stepcount, xyzshift;           /* "Nspace" "Stepcount" "XYZ_Shift_Flag" */
    double amax, anow, astep;                  /* "Amax" "Anow" "Astep" */
    double omep, omeplam, hubble, boxsize;     /* "OmegaMatter0" "OmegaLambda0" "Hubble" "Boxsize(Mpc/h)" */
    double theta, epsilon, rsphere;            /* "Theta" "EPSILON" "Rsphere" */
    size_t header_bytes;                       /* offset of the first particle record */
    size_t record_bytes;                       /* 32 or 40 */
    size_t file_bytes;
} lagpm_snapshot_info;
/* Parses the header up to "#End of Ascii Header" (read_head, Params/header.c:72-105). Keys not
 * listed above are skipped; a listed key whose value does not parse is LAGPM_E_IO. ctx may be
 * NULL (then no message is kept). */
int lagpm_snapshot_read_header(lagpm_ctx* ctx, const char* path, lagpm_snapshot_info* info);
/* Reads the file's DNp dark-matter records (kjhrw.c:98-116) into np = DNp 32-byte records;
 * 40-byte tree records lose their leading list pointer. */
int lagpm_snapshot_read_dm(lagpm_ctx* ctx, const char* path, void* dm_bp, size_t np);
/* Writes dst_path as jwrite does (kjhrw.c:155-220) for a DM-only file: the header lines of
 * src_path with "Anow" (written with 17 digits), "Stepcount"/"Stepnum" and "Particle Type" (0)
 * rewritten, then np 32-byte records, and no trailing SimParameters block: jread freads that
 * block over the parsed header (kjhrw.c:118), so the template's block would restore the
 * template's time and particle type; without it the header stands. np must equal src_path's DNp. */
int lagpm_snapshot_write(lagpm_ctx* ctx, const char* src_path, const char* dst_path, const void* dm_bp, size_t np,
                         double anow, int stepcount);
/* XofP/YofP/ZofP (eunha.h:1212,1228-1232) of np 32-byte records into FP64 arrays (v* may be NULL). */
int lagpm_eunha_decode(const void* dm_bp, size_t np, long nx, double* x, double* y, double* z, double* vx, double* vy,
                       double* vz);
/* The inverse, as lagpm_box_download_eunha writes it (any array may be NULL: that field is kept). */
int lagpm_eunha_encode(void* dm_bp, size_t np, long nx, const double* x, const double* y, const double* z,
                       const double* vx, const double* vy, const double* vz);

/* ---- PM long-range force (pmmain / TSC / psolver / fda4) -------------------------- */
/* Full PM kick: deposit -> r2c -> Green -> c2r -> FDA4 gradient -> TSC interpolation,
 * v += f * F_PM with f = fact2 (KICK), fact2_pull (PULL), fact2_push/GetFact1(a,a+da/2)
 * (PUSH; the caller passes that quotient in fact2_push). Cosmos/pmmain.c:390-527. */
int lagpm_pmmain(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact* ef, int pmstatus);
/* TSC mass deposit with periodic fold, minus 1 per cell (pmmain.c:304). Result stays
 * resident; den (n^3 floats, x fastest) is filled when not NULL. Cosmos/pmmain.h:3-78. */
int lagpm_tsc(lagpm_ctx* ctx, lagpm_box* box, float* den);
/* Raw int64 fixed-point TSC accumulator (n^3, x fastest) and its scale exponent:
 * density = acc * 2^-scale_bits (no -1 shift). For bit-exact parity tests. */
int lagpm_tsc_fixed(lagpm_ctx* ctx, lagpm_box* box, int64_t* acc, int* scale_bits);
/* r2c of the resident density, Green multiply (psolver.c:13-86), c2r. phi (n^3) optional. */
int lagpm_psolver(lagpm_ctx* ctx, lagpm_box* box, float* phi);
/* Green multiply alone on a caller spectrum (n*n*(n/2+1) complex float, x fastest),
 * in place, exactly as Cosmos/psolver.c:49-76 (FP64 factor, FP32 storage). */
int lagpm_green_apply(lagpm_ctx* ctx, int n, double efold, int deconvolve, float* spec_host);
/* FDA4 gradient + TSC interpolation + kick from the resident potential, fda4.c:31-153. */
int lagpm_fda4(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact* ef, int pmstatus);
/* Replace the resident potential with a caller mesh (n^3 floats), for stage tests. */
int lagpm_set_potential(lagpm_ctx* ctx, lagpm_box* box, const float* phi);
/* Download the gradient mesh (n^3 x 4 floats: fx, fy, fz, phi). */
int lagpm_get_gradient(lagpm_ctx* ctx, lagpm_box* box, float* grad4);

/* CorrMeasure (Cosmos/Correlation.c:13-116) of the box's current particles: TSC density minus
 * one, r2c, then |delta_k|^2/n^6 over the modes with kx < n/2, divided by the squared TSC
 * window and binned with linear weights between floor(|k|) and its neighbour (grid units).
 * Outputs (n+1 entries each): pk (bin means), npk (bin weights), cor (xi at r = i cells);
 * sigma8 for a top hat of 8 Mpc/h in a box of side boxsize Mpc/h; slope of log cor as the
 * reference computes it from cor[2] and cor[3]. */
int lagpm_power_spectrum(lagpm_ctx* ctx, lagpm_box* box, double boxsize, double* pk, double* npk, double* cor,
                         double* sigma8, double* slope);
/* lagpm_power_spectrum without any allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (a serving loop that measures P(k) every step, as the
 * reference's pmmain does, or a harness that forbids hipMalloc). The bins live in one static
 * device array per GPU of 2(n+1) doubles for n <= 4096 (65,552 bytes, held until the process
 * ends); a larger n is LAGPM_E_INVAL. The box's FFT plan and execution info must already exist
 * (made by any PM kick or lagpm_power_spectrum on a box of this size and precision on this
 * context); otherwise LAGPM_E_STATE. Calls on one GPU share the array: calls that overlap on one
 * GPU (from several threads or contexts) race and must be serialised by the caller. */
int lagpm_power_spectrum_noalloc(lagpm_ctx* ctx, lagpm_box* box, double boxsize, double* pk, double* npk,
                                 double* cor, double* sigma8, double* slope);

/* ---- short-range tree correction --------------------------------------------------- */
/* Builds FP64 force tables (NSPLINE=4096 entries over r in [0,6)): diff[i][3], slope[i][3]
 * exactly as OST/Model4TreeForce.c:48-91 (law PMFIT), as ForceLaw/treeforce.c:150-186
 * (law TREEFORCE) or for the erfc law. Host only. */
#define LAGPM_NSPLINE 4096
int lagpm_force_spline(const lagpm_params* p, double* diff /*[4096*3]*/, double* slope /*[4096*3]*/);
/* Tree correction kick v += fact2 * a_short (GOTPM: the tree uses EvolFactArr[0].fact2,
 * Cosmos/treecorrection.c:227). Sorts, builds the octree, walks. */
int lagpm_treecorrection(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact* ef);
int lagpm_tree_stats_get(const lagpm_box* box, lagpm_tree_stats* out);
/* Short-range acceleration of selected particles only. No counterpart in the reference: it states
 * force accuracy where every particle's exact pair sum is out of reach (halo cores of 10^5..10^6
 * particles). Builds the box's octree as lagpm_treecorrection does, then walks each listed
 * particle as a group of one, so the opening test is the reference's per-particle test
 * (OST/ost.ver2.c:296-306), and writes its acceleration per unit kick factor to acc[3t..3t+2].
 * ids: caller indices (the upload order), distinct, in [0, np). Velocities are not changed.
 * Allocates scratch for the call. */
int lagpm_tree_accel_sample(lagpm_ctx* ctx, lagpm_box* box, const int64_t* ids, size_t count, double* acc);
/* The octree of the box's last tree build (Make_Tree3 / divide_node3, OST/ost.ver2.c:36-144), for
 * tests and inspection; it describes the positions the box had at that build. Nodes are in
 * pre-order, *nnodes of them (LAGPM_E_INVAL when cap is smaller; *nnodes is still set).
 *   info[4v..4v+3]: first particle, particle count, number of children, level. A node's particles
 *                   are a run of the box's current particle order; order[i] (np entries) is the
 *                   caller index of the particle at position i of that order.
 *   com[4v..4v+3]:  centre of mass (cells; positions wrapped into [0, n)) and mass.
 *   quad[8v..8v+7]: Qxx, Qyy, Qzz, Qxy, Qxz, Qyz about the centre of mass (mass x cells^2), their
 *                   trace, and rmax (cells): the largest particle distance from the centre of mass
 *                   for leaves and nodes of up to 128 particles, an upper bound of it above.
 * Any of info, com, quad, order may be NULL. */
int lagpm_tree_nodes_get(lagpm_ctx* ctx, lagpm_box* box, size_t cap, size_t* nnodes, int32_t* info, double* com,
                         double* quad, uint32_t* order);
/* Tree tuning of a box: nleaf (a node with more particles is split, 4..256) and the target
 * group size (8..128; one wavefront walks one group, with 1 or 3 targets per lane). Defaults
 * 16 and 64. */
int lagpm_box_set_tuning(lagpm_box* box, int nleaf, int group);
int lagpm_box_get_tuning(const lagpm_box* box, int* nleaf, int* group);
/* How target groups are formed. TREE (default): a tree node of at most `group` particles whose
 * parent holds more, consecutive small siblings merged up to `group`. RUNS: each chaining cell's
 * particles, in Morton order, cut into runs of `group` (the last shorter), so all groups but one
 * per cell fill a wavefront. PACKED: tree groups, but the small children of a parent are packed
 * into runs of exactly `group` consecutive particles (a run may end inside a child), so all of a
 * parent's runs but the last fill a wavefront and each stays inside its parent. Either way a group
 * stays inside one chaining cell and the opening test is taken from the group's bounding box; at
 * theta > 0 the forces differ slightly. */
#define LAGPM_GROUPS_TREE 0
#define LAGPM_GROUPS_RUNS 1
#define LAGPM_GROUPS_PACKED 2
int lagpm_box_set_group_mode(lagpm_box* box, int mode);
int lagpm_box_get_group_mode(const lagpm_box* box);
/* Times the tree build + walk of the box's current particles for a set of (nleaf, group, group
 * mode) candidates (the walk adds nothing to the velocities: factor 0), keeps the fastest, and
 * returns nleaf and group (the mode: lagpm_box_get_group_mode). Forces at theta > 0 depend
 * slightly on the tree, so tune once per box, before stepping, and keep the choice. */
int lagpm_box_autotune(lagpm_ctx* ctx, lagpm_box* box, int* nleaf, int* group);
/* Waits for the box's last recorded phase and fills the timings. */
int lagpm_box_timings(lagpm_ctx* ctx, lagpm_box* box, lagpm_timings* out);

/* ---- integrator pieces ----------------------------------------------------------- */
/* Kick factors in FP64 quadrature (Cosmos/CosmosEvolFactor.c:230-256). */
int lagpm_evolfactor(const lagpm_params* p, double anow, double astep, lagpm_evolfact* out);
/* v *= fact1 (Cosmos/globalT.c:102-110). */
int lagpm_hubble_drag(lagpm_ctx* ctx, lagpm_box* box, double fact1);
/* x += pfact * v, wrapped into [0,n) (Cosmos/stepevolv.c:19-113). */
int lagpm_drift(lagpm_ctx* ctx, lagpm_box* box, double pfact);

/* ---- whole steps ------------------------------------------------------------------ */
/* One ordinary GOTPM global step of a box with PMSTATUS = KICK (the loop body of
 * Cosmos/globalT.c:221-306: PM_PULL_AND_PUSH :114-134, then TreeAndPreFoFmain :38-95, then
 * onestepforwardposition): v *= fact1; v += fact2 * (F_PM + F_tree); x += pfact * v, with the
 * factors of GetEvolFactor(anow, astep) (Cosmos/CosmosEvolFactor.c:230-256) from the box's own
 * cosmology. Returns when the step is complete; tree statistics and phase timings are filled.
 * Node arrays too small for the tree are grown and the tree kick and drift redone inside the call.
 * A tree walk whose bounded stack overflows returns LAGPM_E_OVERFLOW and leaves the box after the
 * drag and PM kick: no tree kick (FP32; the FP64 walk has kicked every other particle) and no
 * drift, so the positions are those the step started from. */
int lagpm_kdk_step(lagpm_ctx* ctx, lagpm_box* box, double anow, double astep);
/* The same step for nbox independent boxes (box i at anow[i], astep[i]); no counterpart in
 * the reference, which evolves one box per MPI job. The boxes' PM phases and tree builds run
 * concurrently on a pool of streams, and the tree walks of all FP32 boxes with the same force
 * law run as one persistent launch. Each box's result equals its lagpm_kdk_step result. */
int lagpm_batch_step(lagpm_ctx* ctx, lagpm_box* const* boxes, int nbox, const double* anow, const double* astep);
/* Whole steps capture each box's step (up to the tree walk) in a hipGraph on first use and
 * replay it (default on); off: every kernel is launched each step. Results are identical. */
int lagpm_set_step_graphs(lagpm_ctx* ctx, int enable);
/* FP32 walks also count the evaluated particle pairs that are within the cutoff
 * (lagpm_tree_stats.n_pp_within): the useful fraction of the interaction lists. Costs a compare
 * and an add per pair; off by default. No counterpart in the reference. */
int lagpm_set_walk_census(lagpm_ctx* ctx, int enable);
/* Batch steps of FP32 boxes sharing one force law run as `stages` groups of consecutive boxes
 * (default 5, reserve_cu 32; at most one group per two boxes), whose particle shares grow by 1.5
 * from one group to the next: the first group's front, which nothing overlaps, is short. A group's
 * fused walk leaves reserve_cu CUs free and runs beside the next group's front halves (PM and
 * tree build). stages = 1 queues every front, then one walk. Results are identical either way.
 * No counterpart in the reference. */
int lagpm_set_batch_pipeline(lagpm_ctx* ctx, int stages, int reserve_cu);

#ifdef __cplusplus
}
#endif
#endif /* LAGPM_H */
