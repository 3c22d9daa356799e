// api.hip — the C ABI of include/lagpm.h: contexts, boxes, uploads, the drop-in entry points,
// force tables, kick factors, drift and drag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <cstdlib>
#include <vector>

#include "lagpm_internal.h"

int lagpm_fail(lagpm_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

namespace {

std::once_flag g_rocfft_once;

__global__ void wrap_kernel(double* __restrict__ a, size_t np, double L) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double v = fmod(a[i], L);
    if (v < 0) v += L;
    if (v >= L) v = 0.0;
    a[i] = v;
}
__global__ void iota_kernel(uint32_t* __restrict__ a, size_t np) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) a[i] = (uint32_t)i;
}
__global__ void scatter_kernel(const double* __restrict__ src, const uint32_t* __restrict__ id, size_t np,
                               double* __restrict__ dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) dst[id[i]] = src[i];
}
// kick factors live in device memory (box->dfac) so that a captured step reads this step's values
__global__ void set_factors_kernel(double* __restrict__ dfac, double f1, double f2, double pf, int mask) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (mask & 1) dfac[0] = f1;
    if (mask & 2) dfac[1] = f2;
    if (mask & 4) dfac[2] = pf;
}
__global__ void scale3_kernel(double* __restrict__ a, double* __restrict__ b, double* __restrict__ c, size_t np,
                              const double* __restrict__ fac) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double f = fac[0];
    a[i] *= f;
    b[i] *= f;
    c[i] *= f;
}
// skip (optional): a box whose tree build overflowed got no tree kick yet, so it must not
// drift until the host has redone the kick
// Drift x += pfact v (wrapped into [0, L)). With tv, the tree kick of the step is first added to
// the velocities (v += tv, after the PM kick, as the reference adds them).
__global__ void drift_kernel(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                             double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                             const double* __restrict__ tvx, const double* __restrict__ tvy,
                             const double* __restrict__ tvz, size_t np, const double* __restrict__ fac, double L,
                             const uint32_t* __restrict__ dsz) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // guarded drift of a step: no drift when the tree build overflowed (the step is redone) or a
    // walk stack overflowed (the step returns the error)
    if (i >= np || (dsz && (dsz[DSZ_OVERFLOW] | dsz[DSZ_WALKOVF]))) return;
    const double pfact = fac[2];
    double u = vx[i], w = vy[i], q = vz[i];
    if (tvx) {
        u += tvx[i]; w += tvy[i]; q += tvz[i];
        vx[i] = u; vy[i] = w; vz[i] = q;
    }
    double p[3] = {x[i] + pfact * u, y[i] + pfact * w, z[i] + pfact * q};
    for (int a = 0; a < 3; a++) {
        double v = fmod(p[a], L);
        if (v < 0) v += L;
        if (v >= L) v = 0.0;
        p[a] = v;
    }
    x[i] = p[0];
    y[i] = p[1];
    z[i] = p[2];
}

// v += tv: a tree kick computed on its own (see the walk) applied to the velocities
__global__ void add3_kernel(double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                            const double* __restrict__ tvx, const double* __restrict__ tvy, const double* __restrict__ tvz,
                            size_t np) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    vx[i] += tvx[i];
    vy[i] += tvy[i];
    vz[i] += tvz[i];
}

unsigned nblocks(size_t n, int t) { return (unsigned)std::max<size_t>(1, (n + t - 1) / t); }

template <typename T>
int dalloc(lagpm_ctx* ctx, lagpm_box* box, T** p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return lagpm_fail(ctx, LAGPM_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    box->bytes += bytes;
    return LAGPM_OK;
}
template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int fixed_scale_bits(double total_mass) {
    int e = 0;
    if (total_mass > 0) e = (int)std::ceil(std::log2(total_mass));
    int s = 61 - e;
    return std::min(40, std::max(8, s));
}

void free_box(lagpm_box* b) {
    dfree(b->x); dfree(b->y); dfree(b->z); dfree(b->vx); dfree(b->vy); dfree(b->vz); dfree(b->id);
    dfree(b->key); dfree(b->key2); dfree(b->perm); dfree(b->perm2);
    for (int a = 0; a < 6; a++) dfree(b->alt[a]);
    dfree(b->id_alt); dfree(b->dfac);
    lagpm_box_fft_release(b);
    lagpm_box_graphs_release(b);
    dfree(b->acc); dfree(b->dep_start); dfree(b->rho); dfree(b->grad);
    b->spec = b->phi = nullptr;  // the same padded mesh as rho
    dfree(b->pp); dfree(b->ncount); dfree(b->noff); dfree(b->nd_info); dfree(b->nd_parent); dfree(b->nd_child);
    dfree(b->nd_com); dfree(b->nd_q0); dfree(b->nd_q1); dfree(b->nd_d); dfree(b->cellroot); dfree(b->gflag);
    dfree(b->w_link); dfree(b->w_geo); dfree(b->w_q0); dfree(b->w_q1); dfree(b->w_rfc); dfree(b->w_cellroot);
    dfree(b->dep_scan_tmp); dfree(b->dep_blk);
    dfree(b->goff); dfree(b->groups); dfree(b->gbox); dfree(b->pgroup); dfree(b->counters); dfree(b->dsz);
    dfree(b->scan_tmp);
}

int check_params(lagpm_ctx* ctx, const lagpm_params* p) {
    if (!p) return lagpm_fail(ctx, LAGPM_E_INVAL, "params is NULL");
    if (p->n < 8 || p->n > 4096) return lagpm_fail(ctx, LAGPM_E_INVAL, "n must be in [8, 4096]");
    if (p->law != LAGPM_LAW_PMFIT && p->law != LAGPM_LAW_ERFC && p->law != LAGPM_LAW_TREEFORCE)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "unknown force law");
    if (p->precision != LAGPM_PREC_FP32 && p->precision != LAGPM_PREC_FP64)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "unknown precision");
    if (!(p->mass > 0)) return lagpm_fail(ctx, LAGPM_E_INVAL, "mass must be positive");
    if (!(p->efold >= 0)) return lagpm_fail(ctx, LAGPM_E_INVAL, "efold must be >= 0");
    if (!(p->rsphere > 0 && p->rsphere < 6.0)) return lagpm_fail(ctx, LAGPM_E_INVAL, "rsphere must be in (0, 6) (force table range)");
    if (!(p->theta >= 0)) return lagpm_fail(ctx, LAGPM_E_INVAL, "theta must be >= 0");
    if (!(p->epsilon >= 0)) return lagpm_fail(ctx, LAGPM_E_INVAL, "epsilon must be >= 0");
    return LAGPM_OK;
}

// ---- FP64 quadrature for the kick factors: composite 16-point Gauss-Legendre ----
const double GL_X[8] = {0.0950125098376374, 0.2816035507792589, 0.4580167776572274, 0.6178762444026438,
                        0.7554044083550030, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499};
const double GL_W[8] = {0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                        0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541};
struct cosmo { double om, ol, amax; };
double hofez(const cosmo& c, double a) {
    double zp1 = c.amax / a;
    return std::sqrt(zp1 * zp1 * zp1 * c.om + zp1 * zp1 * (1 - c.om - c.ol) + c.ol);
}
double apf(const cosmo& c, double a) {
    double zp1 = c.amax / a, h = hofez(c, a);
    return std::sqrt(8.0 * M_PI / 3.0 / (c.amax * c.om) / (zp1 * zp1) * h * h);
}
double integ(const cosmo& c, double a1, double a2) {
    const int nsub = 4;  // 8-point Gauss-Legendre on 4 pieces: the integrand is smooth over a step
    double s = 0, hstep = (a2 - a1) / nsub;
    for (int k = 0; k < nsub; k++) {
        double lo = a1 + k * hstep, mid = lo + 0.5 * hstep, half = 0.5 * hstep;
        for (int q = 0; q < 8; q++) {
            double u1 = mid - half * GL_X[q], u2 = mid + half * GL_X[q];
            s += GL_W[q] * half * (1.0 / (u1 * apf(c, u1)) + 1.0 / (u2 * apf(c, u2)));
        }
    }
    return s;
}
double fact1_of(const cosmo& c, double a1, double a2) {
    return (a1 * a1 * a1 * hofez(c, a1)) / (a2 * a2 * a2 * hofez(c, a2));
}
double fact2_of(const cosmo& c, int n, double a1, double a2) {
    return 1.0 / (a2 * a2 * apf(c, a2)) * integ(c, a1, a2) * n;
}

// ---- force-law tables (FP64), OST/Model4TreeForce.c:48-116 ----
const double PMFITPAR[8] = {1.0 / 256.0 / 256.0, 0.78224e+00, 0.37971e-06, -0.60338e+00,
                            -0.34419e-07, -0.39741e+01, -0.10607e+01, -0.38145e+00};
double law_pmfit(int n, double eps, double r) {
    const double* a = PMFITPAR;
    double ch = std::cosh(a[1] * r);
    double newton = r / std::pow(r * r + eps * eps, 1.5);
    double fit = 1.0 / (r * r) * std::tanh(a[1] * r) - a[1] / (ch * ch * r) - 2.0 * a[2] / a[0] * r * std::exp(a[3] * r * r) -
                 2.0 * a[2] * a[3] / a[0] * r * r * r * std::exp(a[3] * r * r) +
                 a[4] / a[0] * (1.0 + a[5] * r * r + a[6] * r * r * r * r) * std::exp(a[7] * r * r);
    return (newton - fit) / ((double)n * n);
}
double law_erfc(int n, double eps, double rs, double r) {
    double newton = r / std::pow(r * r + eps * eps, 1.5);
    double u = r / (2.0 * rs);
    double flong = (std::erf(u) - (r / (rs * std::sqrt(M_PI))) * std::exp(-u * u)) / (r * r);
    return (newton - flong) / ((double)n * n);
}

}  // namespace

// LAW_PMFIT: i_force_spline of OST/Model4TreeForce.c:48-91 (softening GRV_EPSILON, derivative
// step 0.1 bin, entry 0 a copy of entry 1). LAW_TREEFORCE: i_force_spline(nx, Epsilon) of
// ForceLaw/treeforce.c:150-186, the same fitted law with the softening fixed at the EPSILON macro
// (0.1, OST/ost.h:6; its pmfit ignores the Epsilon argument, :134-147), derivative step 0.2 bin
// (:159), and entry 0 left zero with the slope up to entry 1 (:153,170-174). LAW_ERFC: the erfc
// complement of the Gaussian filter, built like LAW_PMFIT.
void lagpm_build_tables(const lagpm_params* p, double* diff, double* slope) {
    const int NS = LAGPM_NSPLINE;
    const bool tf = p->law == LAGPM_LAW_TREEFORCE;
    const double ran = 6.0 / (double)NS, dr = (tf ? 0.2 : 0.1) * ran;
    const double rs = 0.5 * p->efold;
    auto law = [&](double r) {
        if (p->law == LAGPM_LAW_PMFIT) return law_pmfit(p->n, p->epsilon, r);
        if (tf) return law_pmfit(p->n, 0.1, r);
        return law_erfc(p->n, p->epsilon, rs, r);
    };
    std::memset(diff, 0, 3 * NS * sizeof(double));
    std::memset(slope, 0, 3 * NS * sizeof(double));
    for (int i = 1; i < NS; i++) {
        double r = i * ran;
        double fp = (law(r + dr) - law(r - dr)) / (dr + dr);
        double fpp = (law(r + 2 * dr) - 2 * law(r) + law(r - 2 * dr)) / (4 * dr * dr);
        double fr = law(r);
        diff[3 * i + 0] = -(fr / r);
        diff[3 * i + 1] = -(0.5 * (fp * r - fr) / r / r / r);
        diff[3 * i + 2] = -(0.5 * (fpp * r * r - 3 * fp * r + 3 * fr) / r / r / r / r / r);
    }
    for (int i = tf ? 0 : 1; i < NS - 1; i++)
        for (int c = 0; c < 3; c++) slope[3 * i + c] = (diff[3 * (i + 1) + c] - diff[3 * i + c]) / ran;
    for (int c = 0; c < 3; c++) {
        if (!tf) {
            slope[c] = slope[3 + c];
            diff[c] = diff[3 + c];
        }
        diff[3 * (NS - 1) + c] = 0;
        slope[3 * (NS - 1) + c] = 0;
    }
}

int lagpm_dalloc_bytes(lagpm_ctx* ctx, void** p, size_t bytes) {
    HIPCHK(ctx, hipMalloc(p, std::max<size_t>(bytes, 1)));
    return LAGPM_OK;
}

// g(s) = fr(r)/r at s = r^2 for n = 1 (the short-range force per unit separation); in terms of
// s the reference's table terms are f0 = -g, f1 = -(fp r - fr)/(2 r^3) = -dg/ds and
// f2 = -(fpp r^2 - 3 fp r + 3 fr)/(2 r^5) = -2 d2g/ds2 (OST/Model4TreeForce.c:48-91).
static double erfc_g(double eps, double rs, double s) {
    const double newton = 1.0 / std::pow(s + eps * eps, 1.5);
    const double u2 = s / (4.0 * rs * rs);
    double flong;
    if (u2 < 1e-6) {  // series of (erf(u) - 2u/sqrt(pi) exp(-u^2)) / r^3 in u^2 (also for s slightly < 0)
        flong = (4.0 / (3.0 * std::sqrt(M_PI))) / (8.0 * rs * rs * rs) * (1.0 - 0.6 * u2 + (3.0 / 14.0) * u2 * u2);
    } else {
        const double r = std::sqrt(s), u = std::sqrt(u2);
        flong = (std::erf(u) - (r / (rs * std::sqrt(M_PI))) * std::exp(-u2)) / (s * r);
    }
    return newton - flong;
}

// Bins of the erfc law, uniform in log(lam r^2) with lam = 32 / rc2: bin b holds the line
// c + s * r2 through the law at its node and the next node (bin 0 from r2 = 0). The node
// lam r2 = 32 is the cutoff; bins from there on hold zeros.
static void erfc_log_tables(double eps, double rs, double rc2, double lam, std::vector<double>& t /* 6 per bin */) {
    const int NB = LAGPM_LOG_BINS;
    auto node = [](int b) -> double {
        uint32_t bits = (uint32_t)(b + (int)LAGPM_LOG_BASE) << LAGPM_LOG_SHIFT;
        float f;
        std::memcpy(&f, &bits, 4);
        return (double)f;
    };
    auto terms = [&](double s, double out[3]) {
        const double h = 1e-4 * std::max(s, eps * eps);
        const double g0 = erfc_g(eps, rs, s), gp = erfc_g(eps, rs, s + h), gm = erfc_g(eps, rs, s - h);
        out[0] = -g0;
        out[1] = -(gp - gm) / (2 * h);
        out[2] = -2.0 * (gp - 2 * g0 + gm) / (h * h);
    };
    t.assign(6 * (size_t)NB, 0.0);
    for (int b = 0; b < NB; b++) {
        if (node(b) >= 32.0) break;  // lam r2 >= 32 is r2 >= rc2: zeros
        // bin 0 also serves every r^2 below its node: its line runs from r^2 = 0 to the next node
        const double s0 = b == 0 ? 0.0 : node(b) / lam, s1 = node(b + 1) / lam;
        double a[3], c[3];
        terms(s0, a);
        terms(s1, c);
        for (int k = 0; k < 3; k++) {
            const double slope = (c[k] - a[k]) / (s1 - s0);
            t[6 * (size_t)b + 2 * k] = a[k] - slope * s0;  // intercept at r2 = 0
            t[6 * (size_t)b + 2 * k + 1] = slope;
        }
    }
    (void)rc2;
}

// Device force tables of a law, built once per context for n = 1 (the law scales as 1/n^2;
// the walk multiplies by the box's mass / n^2 with the kick factor).
int lagpm_get_tables(lagpm_ctx* ctx, const lagpm_params* p, const law_tables** out) {
    const auto key = std::make_tuple(p->law, p->epsilon, p->efold, p->rsphere);
    auto it = ctx->tables.find(key);
    if (it != ctx->tables.end()) {
        *out = &it->second;
        return LAGPM_OK;
    }
    law_tables lt;
    std::vector<double> t;  // 6 per bin: d0 s0 d1 s1 d2 s2
    if (p->law == LAGPM_LAW_ERFC) {
        lt.logr2 = 1;
        lt.nbins = LAGPM_LOG_BINS;
        const double rc2 = p->rsphere * p->rsphere;
        lt.lam = (float)(32.0 / rc2);
        lt.rc2_eff = rc2;
        erfc_log_tables(p->epsilon, 0.5 * p->efold, rc2, (double)lt.lam, t);
    } else {
        const int NS = LAGPM_NSPLINE;
        lagpm_params q = *p;
        q.n = 1;
        std::vector<double> diff(3 * NS), slope(3 * NS);
        lagpm_build_tables(&q, diff.data(), slope.data());
        lt.logr2 = 0;
        lt.nbins = NS;
        lt.rc2_eff = p->rsphere * p->rsphere;
        t.resize(6 * (size_t)NS);
        for (int i = 0; i < NS; i++)
            for (int c = 0; c < 3; c++) {
                t[6 * i + 2 * c] = diff[3 * i + c];
                t[6 * i + 2 * c + 1] = slope[3 * i + c];
            }
    }
    const int NB = lt.nbins;
    std::vector<float2> t0(NB);
    std::vector<float4> t12(NB);
    // The FP32 walk evaluates the log tables in coordinates scaled by sqrt(lam) (the tree's pp
    // copy), where its variable u = lam r^2 is the plain squared distance. With d' = sqrt(lam) d and
    // Q' = lam Q the kick times sqrt(lam) is d' (M f0 + tr(Q') f1 / lam + d'Q'd' f2 / lam^2)
    // + 2 (f1 / lam) Q'd', so the FP32 lines are f0, f1 / lam and f2 / lam^2 as lines in u.
    // The f1 line is stored doubled (exact: a power of two), so the walk's 2 f1 Q'd' needs no
    // multiply; the walk's node copy holds tr(Q')/2 to match, and tr(Q') f1 is unchanged bit for bit.
    // The FP64 walk keeps the plain lines in r^2 (tabd).
    const double L = lt.logr2 ? (double)lt.lam : 1.0;
    for (int i = 0; i < NB; i++) {
        t0[i] = make_float2((float)t[6 * i], (float)(t[6 * i + 1] / L));
        t12[i] = make_float4(2.0f * (float)(t[6 * i + 2] / L), 2.0f * (float)(t[6 * i + 3] / (L * L)),
                             (float)(t[6 * i + 4] / (L * L)), (float)(t[6 * i + 5] / (L * L * L)));
    }
    RCHK(lagpm_dalloc_bytes(ctx, (void**)&lt.tab0, NB * sizeof(float2)));
    RCHK(lagpm_dalloc_bytes(ctx, (void**)&lt.tab12, NB * sizeof(float4)));
    RCHK(lagpm_dalloc_bytes(ctx, (void**)&lt.tabd, 6 * NB * sizeof(double)));
    HIPCHK(ctx, hipMemcpyAsync(lt.tab0, t0.data(), NB * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(lt.tab12, t12.data(), NB * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(lt.tabd, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    auto res = ctx->tables.emplace(key, lt);
    *out = &res.first->second;
    return LAGPM_OK;
}

int lagpm_d2h(lagpm_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t s) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    const hipError_t w = hipStreamSynchronize(s);
    if (e != hipSuccess) return lagpm_fail(ctx, LAGPM_E_HIP, std::string("device-to-host copy: ") + hipGetErrorString(e));
    if (w != hipSuccess) return lagpm_fail(ctx, LAGPM_E_HIP, std::string("device-to-host copy (wait): ") + hipGetErrorString(w));
    return LAGPM_OK;
}

extern "C" {

int lagpm_abi_version(void) { return LAGPM_ABI_VERSION; }


int lagpm_create(lagpm_ctx** out, int device) {
    if (!out) return LAGPM_E_INVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LAGPM_E_HIP;
    if (device < 0 || device >= ndev) return LAGPM_E_INVAL;
    if (hipSetDevice(device) != hipSuccess) return LAGPM_E_HIP;
    lagpm_ctx* ctx = new lagpm_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return LAGPM_E_HIP;
    }
    ctx->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return LAGPM_E_HIP;
    }
    for (int i = 0; i < 4; i++) (void)hipEventCreate(&ctx->ev[i]);
    ctx->walk_stack_entries = lagpm_walk_stack_entries(ctx->num_cu);
    if (hipMalloc((void**)&ctx->walk_stack, ctx->walk_stack_entries * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&ctx->walk_ticket, LAGPM_WALK_XCDS * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        lagpm_destroy(ctx);
        return LAGPM_E_NOMEM;
    }
    // The library never writes the process environment: setenv() is not thread-safe, and by
    // now the HIP runtime has threads of its own (a caller that wants rocFFT's run-time kernel
    // cache in a chosen file sets ROCFFT_RTC_CACHE_PATH before the first context; the Python
    // binding does so before it loads this library).
    std::call_once(g_rocfft_once, []() { rocfft_setup(); });
    if (lagpm_pk_scratch_init(ctx) != LAGPM_OK) {
        lagpm_destroy(ctx);
        return LAGPM_E_HIP;
    }
    *out = ctx;
    return LAGPM_OK;
}

int lagpm_destroy(lagpm_ctx* ctx) {
    if (!ctx) return LAGPM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    lagpm_fft_release(ctx);
    for (auto& kv : ctx->tables) {
        dfree(kv.second.tab0);
        dfree(kv.second.tab12);
        dfree(kv.second.tabd);
    }
    dfree(ctx->walk_stack);
    dfree(ctx->walk_ticket);
    dfree(ctx->walk_boxes);
    dfree(ctx->walk_prefix);
    if (ctx->readback) (void)hipHostFree(ctx->readback);
    if (ctx->readback1) (void)hipHostFree(ctx->readback1);
    for (int i = 0; i < 4; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    for (hipStream_t s : ctx->pool) (void)hipStreamDestroy(s);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LAGPM_OK;
}

const char* lagpm_last_error(const lagpm_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }
void* lagpm_stream(lagpm_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }
int lagpm_synchronize(lagpm_ctx* ctx) {
    if (!ctx) return LAGPM_E_INVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return LAGPM_OK;
}

size_t lagpm_estimate_box_bytes(const lagpm_params* p, size_t np) {
    if (!p || p->n < 1) return 0;
    const size_t n = (size_t)p->n, m = n * n * n, K = (n / 2 + 1) * n * n;
    const bool f64 = p->precision == LAGPM_PREC_FP64;
    const size_t real = f64 ? 8 : 4;
    const int L = lagpm_choose_cell_level(p->n, p->rsphere);
    const size_t mcell = (size_t)1 << L;
    size_t b = np * (12 * 8 + 4 * 4 + 2 * 8 + 2 * 4);  // two sets of x..vz, id/id_alt/perm/perm2, keys, node counts
    b += np * (f64 ? 4 : 16);                             // pgroup / pp
    b += m * 8 + K * 2 * real + 4 * m * real;              // acc, padded FFT mesh, grad
    b += (f64 ? 4 : 8) * mcell * mcell * mcell + 4 * 8 + DSZ_COUNT * 4;  // cell roots, counters, sizes
    b += 4 * (lagpm_deposit_chunk_cap(p->n, np) + 1);     // deposit chunk starts
    b += 12 * (lagpm_deposit_blocks(p->n) + 1) + lagpm_scan_tmp_bytes(lagpm_deposit_blocks(p->n));  // chunk list
    const size_t cap = lagpm_tree_node_cap_guess(np, (int)mcell), gcap = cap + np / 64 + 1;
    b += cap * (16 + 4 + 32 + 4 + 4 + (f64 ? 96 : 48 + 68)) + gcap * (8 + (f64 ? 48 : 0));  // (+ the FP32 walk's copy)
    b += std::max(lagpm_scan_tmp_bytes(cap), std::max(lagpm_scan_tmp_bytes(np), lagpm_sort_tmp_bytes(np)));
    return b;
}

int lagpm_box_create(lagpm_ctx* ctx, const lagpm_params* p, size_t np, lagpm_box** out) {
    if (!ctx || !out) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    *out = nullptr;
    RCHK(check_params(ctx, p));
    if (np > (size_t)LAGPM_MAX_INDEX) return lagpm_fail(ctx, LAGPM_E_INVAL, "np must be at most 2^27");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    lagpm_box* box = new lagpm_box();
    box->p = *p;
    box->np = np;
    const size_t n = (size_t)p->n, m = n * n * n, K = (n / 2 + 1) * n * n;
    const bool f64 = p->precision == LAGPM_PREC_FP64;
    const size_t real = f64 ? 8 : 4;
    box->level_cell = lagpm_choose_cell_level(p->n, p->rsphere);
    box->mcell = 1 << box->level_cell;
    box->scale_bits = fixed_scale_bits((double)np * p->mass);
    int rc = LAGPM_OK;
#define ALLOC(ptr, cnt)                                                 \
    do {                                                                \
        if (rc == LAGPM_OK) rc = dalloc(ctx, box, &(ptr), (cnt));       \
    } while (0)
    ALLOC(box->x, np); ALLOC(box->y, np); ALLOC(box->z, np);
    ALLOC(box->vx, np); ALLOC(box->vy, np); ALLOC(box->vz, np);
    ALLOC(box->id, np); ALLOC(box->key, np); ALLOC(box->key2, np); ALLOC(box->perm, np); ALLOC(box->perm2, np);
    for (int a = 0; a < 6; a++) ALLOC(box->alt[a], np);
    ALLOC(box->id_alt, np); ALLOC(box->dfac, 4);
    ALLOC(box->acc, m);
    // one in-place FFT mesh (padded rows of 2(n/2+1)): density, then spectrum, then potential
    if (rc == LAGPM_OK) rc = dalloc(ctx, box, (char**)&box->rho, K * 2 * real);
    box->spec = box->phi = box->rho;
    if (rc == LAGPM_OK) rc = dalloc(ctx, box, (char**)&box->grad, m * 4 * real);
    ALLOC(box->ncount, np); ALLOC(box->noff, np);
    ALLOC(box->cellroot, (size_t)box->mcell * box->mcell * box->mcell);
    if (!f64) ALLOC(box->w_cellroot, (size_t)box->mcell * box->mcell * box->mcell);
    ALLOC(box->counters, 4);
    ALLOC(box->dsz, DSZ_COUNT);
    const size_t nblk = lagpm_deposit_blocks(p->n);
    if (rc == LAGPM_OK) rc = dalloc(ctx, box, (char**)&box->dep_scan_tmp, lagpm_scan_tmp_bytes(nblk));
    ALLOC(box->dep_blk, 3 * (nblk + 1));
    box->dep_cap = lagpm_deposit_chunk_cap(p->n, np) + 1;
    ALLOC(box->dep_start, box->dep_cap);
    if (f64) ALLOC(box->pgroup, np);
    else ALLOC(box->pp, np);
#undef ALLOC
    if (rc == LAGPM_OK) {
        // node arrays and scan scratch (grown later if an input needs more nodes)
        rc = lagpm_tree_alloc(ctx, box, lagpm_tree_node_cap_guess(np, box->mcell));
        if (rc == LAGPM_OK) {
            box->bytes += box->scan_tmp_bytes + box->node_cap * (16 + 4 + 32 + 4 + 4 + (f64 ? 96 : 48)) +
                          box->group_cap * (8 + (f64 ? 48 : 0));
            // the node arrays may grow later (dense input); the estimate is what admission uses
            box->bytes = std::max(box->bytes, lagpm_estimate_box_bytes(p, np));
        }
    }
    if (rc == LAGPM_OK) rc = lagpm_get_tables(ctx, p, &box->tab);
    if (rc != LAGPM_OK) {
        free_box(box);
        delete box;
        return rc;
    }
    box->st = ctx->stream;
    for (int i = 0; i < 15; i++) (void)hipEventCreate(&box->ev[i]);
    (void)hipEventCreateWithFlags(&box->evfork, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&box->evsort, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&box->evvel, hipEventDisableTiming);
    (void)hipMemsetAsync(box->vx, 0, np * sizeof(double), ctx->stream);
    (void)hipMemsetAsync(box->vy, 0, np * sizeof(double), ctx->stream);
    (void)hipMemsetAsync(box->vz, 0, np * sizeof(double), ctx->stream);
    (void)hipMemsetAsync(box->dsz, 0, DSZ_COUNT * sizeof(uint32_t), ctx->stream);
    *out = box;
    return LAGPM_OK;
}

int lagpm_box_destroy(lagpm_ctx* ctx, lagpm_box* box) {
    if (!box) return LAGPM_OK;
    if (ctx) (void)hipStreamSynchronize(box->st);
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 15; i++)
        if (box->ev[i]) (void)hipEventDestroy(box->ev[i]);
    if (box->evfork) (void)hipEventDestroy(box->evfork);
    if (box->evsort) (void)hipEventDestroy(box->evsort);
    if (box->evvel) (void)hipEventDestroy(box->evvel);
    free_box(box);
    delete box;
    return LAGPM_OK;
}

int lagpm_box_params(const lagpm_box* box, lagpm_params* out) {
    if (!box || !out) return LAGPM_E_INVAL;
    *out = box->p;
    return LAGPM_OK;
}
int lagpm_box_set_cosmology(lagpm_box* box, double amax, double omep, double omeplam) {
    if (!box || !(amax > 0) || !(omep > 0) || !(omeplam >= 0)) return LAGPM_E_INVAL;
    // the kick factors come from these at every step; nothing on the device depends on them
    box->p.amax = amax;
    box->p.omep = omep;
    box->p.omeplam = omeplam;
    return LAGPM_OK;
}
size_t lagpm_box_np(const lagpm_box* box) { return box ? box->np : 0; }
size_t lagpm_box_device_bytes(const lagpm_box* box) { return box ? box->bytes : 0; }

int lagpm_box_upload(lagpm_ctx* ctx, lagpm_box* box, const double* x, const double* y, const double* z, const double* vx,
                     const double* vy, const double* vz) {
    if (!ctx || !box || (box->np && (!x || !y || !z))) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    const size_t np = box->np, b = np * sizeof(double);
    if (np == 0) {
        box->uploaded = true;
        return LAGPM_OK;
    }
    HIPCHK(ctx, hipMemcpyAsync(box->x, x, b, hipMemcpyHostToDevice, box->st));
    HIPCHK(ctx, hipMemcpyAsync(box->y, y, b, hipMemcpyHostToDevice, box->st));
    HIPCHK(ctx, hipMemcpyAsync(box->z, z, b, hipMemcpyHostToDevice, box->st));
    double* vs[3] = {box->vx, box->vy, box->vz};
    const double* hs[3] = {vx, vy, vz};
    for (int a = 0; a < 3; a++) {
        if (hs[a]) HIPCHK(ctx, hipMemcpyAsync(vs[a], hs[a], b, hipMemcpyHostToDevice, box->st));
        else HIPCHK(ctx, hipMemsetAsync(vs[a], 0, b, box->st));
    }
    double* ps[3] = {box->x, box->y, box->z};
    for (int a = 0; a < 3; a++)
        hipLaunchKernelGGL(wrap_kernel, dim3(nblocks(np, 256)), dim3(256), 0, box->st, ps[a], np, (double)box->p.n);
    hipLaunchKernelGGL(iota_kernel, dim3(nblocks(np, 256)), dim3(256), 0, box->st, box->id, np);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(box->st));
    box->uploaded = true;
    box->sorted = false;
    box->tree_built = false;
    // new particles: the box starts over as a new box would, with a full-depth first tree (the
    // depth limit also sets how many key bits the sort orders, and so the particle order that
    // the sums follow), so a box reused for other particles steps exactly as a fresh one
    if (box->depth != LAGPM_KEY_BITS) {
        box->depth = LAGPM_KEY_BITS;
        lagpm_box_graphs_release(box);  // captured steps hold the old limit
    }
    return LAGPM_OK;
}

int lagpm_box_download(lagpm_ctx* ctx, lagpm_box* box, double* x, double* y, double* z, double* vx, double* vy, double* vz) {
    if (!ctx || !box) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    const size_t np = box->np;
    double* src[6] = {box->x, box->y, box->z, box->vx, box->vy, box->vz};
    double* dst[6] = {x, y, z, vx, vy, vz};
    for (int a = 0; a < 6; a++) {
        if (!dst[a] || np == 0) continue;
        hipLaunchKernelGGL(scatter_kernel, dim3(nblocks(np, 256)), dim3(256), 0, box->st, src[a], box->id, np, box->alt[0]);
        HIPCHK(ctx, hipGetLastError());
        RCHK(lagpm_d2h(ctx, dst[a], box->alt[0], np * sizeof(double), box->st));
    }
    return LAGPM_OK;
}

// GOTPM + XYZDBL particle: u64 index|flag (flag in the top byte), f32 dx,dy,dz, f32 vx,vy,vz.
int lagpm_box_upload_eunha(lagpm_ctx* ctx, lagpm_box* box, const void* dm_bp, size_t np, long nx) {
    if (!ctx || !box || (np && !dm_bp)) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    if (np != box->np) return lagpm_fail(ctx, LAGPM_E_INVAL, "particle count differs from the box");
    if (nx != box->p.n) return lagpm_fail(ctx, LAGPM_E_INVAL, "nx differs from the box mesh size");
    std::vector<double> a(6 * np);
    double* q = a.data();
    if (lagpm_eunha_decode(dm_bp, np, nx, q, q + np, q + 2 * np, q + 3 * np, q + 4 * np, q + 5 * np) != LAGPM_OK)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "bad particle records");
    return lagpm_box_upload(ctx, box, q, q + np, q + 2 * np, q + 3 * np, q + 4 * np, q + 5 * np);
}

int lagpm_box_download_eunha_vel(lagpm_ctx* ctx, lagpm_box* box, void* dm_bp, size_t np) {
    if (!ctx || !box || (np && !dm_bp)) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    if (np != box->np) return lagpm_fail(ctx, LAGPM_E_INVAL, "particle count differs from the box");
    std::vector<double> v(3 * np);
    RCHK(lagpm_box_download(ctx, box, nullptr, nullptr, nullptr, v.data(), v.data() + np, v.data() + 2 * np));
    return lagpm_eunha_encode(dm_bp, np, box->p.n, nullptr, nullptr, nullptr, v.data(), v.data() + np, v.data() + 2 * np);
}

int lagpm_box_download_eunha(lagpm_ctx* ctx, lagpm_box* box, void* dm_bp, size_t np) {
    if (!ctx || !box || (np && !dm_bp)) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    if (np != box->np) return lagpm_fail(ctx, LAGPM_E_INVAL, "particle count differs from the box");
    std::vector<double> a(6 * np);
    double* q = a.data();
    RCHK(lagpm_box_download(ctx, box, q, q + np, q + 2 * np, q + 3 * np, q + 4 * np, q + 5 * np));
    return lagpm_eunha_encode(dm_bp, np, box->p.n, q, q + np, q + 2 * np, q + 3 * np, q + 4 * np, q + 5 * np);
}

static int pm_factor(lagpm_ctx* ctx, const lagpm_evolfact* ef, int pmstatus, double* f) {
    if (!ef) return lagpm_fail(ctx, LAGPM_E_INVAL, "evolfact is NULL");
    switch (pmstatus) {
        case LAGPM_KICK: *f = ef->fact2; break;
        case LAGPM_PULL: *f = ef->fact2_pull; break;
        case LAGPM_PUSH:
            if (ef->fact1_push == 0.0f) return lagpm_fail(ctx, LAGPM_E_INVAL, "fact1_push is zero");
            *f = (double)(ef->fact2_push / ef->fact1_push);  // fda4.c:135-141, GetFact1(a, a+da/2) == fact1_push
            break;
        default: return lagpm_fail(ctx, LAGPM_E_INVAL, "unknown pmstatus");
    }
    return LAGPM_OK;
}

static int need_particles(lagpm_ctx* ctx, lagpm_box* box) {
    if (!ctx || !box) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    if (!box->uploaded) return lagpm_fail(ctx, LAGPM_E_STATE, "no particles uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return LAGPM_OK;
}

// n^3 host values <-> the padded FFT mesh (rows of n values every 2(n/2+1))
static int mesh_copy(lagpm_ctx* ctx, lagpm_box* box, void* host, bool to_host, size_t elem) {
    const size_t n = (size_t)box->p.n, mx = 2 * (n / 2 + 1);
    if (to_host)
        HIPCHK(ctx, hipMemcpy2DAsync(host, n * elem, box->rho, mx * elem, n * elem, n * n, hipMemcpyDeviceToHost, box->st));
    else
        HIPCHK(ctx, hipMemcpy2DAsync(box->rho, mx * elem, host, n * elem, n * elem, n * n, hipMemcpyHostToDevice, box->st));
    return LAGPM_OK;
}

int lagpm_tsc(lagpm_ctx* ctx, lagpm_box* box, float* den) {
    RCHK(need_particles(ctx, box));
    if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
    RCHK(lagpm_pm_deposit(ctx, box));
    if (den) {
        const size_t m = (size_t)box->p.n * box->p.n * box->p.n;
        if (box->p.precision == LAGPM_PREC_FP64) {
            std::vector<double> t(m);
            RCHK(mesh_copy(ctx, box, t.data(), true, 8));
            HIPCHK(ctx, hipStreamSynchronize(box->st));
            for (size_t i = 0; i < m; i++) den[i] = (float)t[i];
        } else {
            RCHK(mesh_copy(ctx, box, den, true, 4));
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(box->st));
    return LAGPM_OK;
}

int lagpm_tsc_fixed(lagpm_ctx* ctx, lagpm_box* box, int64_t* acc, int* scale_bits) {
    RCHK(need_particles(ctx, box));
    if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
    RCHK(lagpm_pm_deposit(ctx, box));
    const size_t m = (size_t)box->p.n * box->p.n * box->p.n;
    if (acc) RCHK(lagpm_d2h(ctx, acc, box->acc, m * 8, box->st));
    if (scale_bits) *scale_bits = box->scale_bits;
    HIPCHK(ctx, hipStreamSynchronize(box->st));
    return LAGPM_OK;
}

int lagpm_psolver(lagpm_ctx* ctx, lagpm_box* box, float* phi) {
    RCHK(need_particles(ctx, box));
    RCHK(lagpm_pm_solve(ctx, box));
    if (phi) {
        const size_t m = (size_t)box->p.n * box->p.n * box->p.n;
        if (box->p.precision == LAGPM_PREC_FP64) {
            std::vector<double> t(m);
            RCHK(mesh_copy(ctx, box, t.data(), true, 8));
            HIPCHK(ctx, hipStreamSynchronize(box->st));
            for (size_t i = 0; i < m; i++) phi[i] = (float)t[i];
        } else {
            RCHK(mesh_copy(ctx, box, phi, true, 4));
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(box->st));
    return LAGPM_OK;
}

int lagpm_green_apply(lagpm_ctx* ctx, int n, double efold, int deconvolve, float* spec_host) {
    if (!ctx || !spec_host || n < 2) return lagpm_fail(ctx, LAGPM_E_INVAL, "bad argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t K = (size_t)(n / 2 + 1) * n * n;
    void* d = nullptr;
    HIPCHK(ctx, hipMalloc(&d, K * 2 * sizeof(float)));
    int rc = LAGPM_OK;
    if (hipMemcpyAsync(d, spec_host, K * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = LAGPM_E_HIP;
    if (rc == LAGPM_OK) rc = lagpm_green_prepare(ctx, ctx->stream, n, efold, deconvolve);
    if (rc == LAGPM_OK) rc = lagpm_green_device(ctx, ctx->stream, n, LAGPM_PREC_FP32, efold, deconvolve, d);
    if (rc == LAGPM_OK && hipMemcpyAsync(spec_host, d, K * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) rc = LAGPM_E_HIP;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (rc == LAGPM_E_HIP) return lagpm_fail(ctx, rc, "green_apply copy failed");
    return rc;
}

int lagpm_fda4(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact* ef, int pmstatus) {
    RCHK(need_particles(ctx, box));
    double f;
    RCHK(pm_factor(ctx, ef, pmstatus, &f));
    RCHK(lagpm_pm_gradient(ctx, box));
    hipLaunchKernelGGL(set_factors_kernel, dim3(1), dim3(64), 0, box->st, box->dfac, 0.0, f, 0.0, 2);
    RCHK(lagpm_pm_interp(ctx, box));
    HIPCHK(ctx, hipStreamSynchronize(box->st));
    return LAGPM_OK;
}

int lagpm_set_potential(lagpm_ctx* ctx, lagpm_box* box, const float* phi) {
    if (!ctx || !box || !phi) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    const size_t m = (size_t)box->p.n * box->p.n * box->p.n;
    if (box->p.precision == LAGPM_PREC_FP64) {
        std::vector<double> t(phi, phi + m);
        RCHK(mesh_copy(ctx, box, t.data(), false, 8));
        HIPCHK(ctx, hipStreamSynchronize(box->st));
    } else {
        RCHK(mesh_copy(ctx, box, (void*)phi, false, 4));
        HIPCHK(ctx, hipStreamSynchronize(box->st));
    }
    return LAGPM_OK;
}

int lagpm_get_gradient(lagpm_ctx* ctx, lagpm_box* box, float* grad4) {
    if (!ctx || !box || !grad4) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    const size_t m = (size_t)box->p.n * box->p.n * box->p.n;
    if (box->p.precision == LAGPM_PREC_FP64) {
        std::vector<double> t(4 * m);
        RCHK(lagpm_d2h(ctx, t.data(), box->grad, 4 * m * 8, box->st));
        for (size_t i = 0; i < 4 * m; i++) grad4[i] = (float)t[i];
    } else {
        RCHK(lagpm_d2h(ctx, grad4, box->grad, 4 * m * 4, box->st));
    }
    return LAGPM_OK;
}

int lagpm_pmmain(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact* ef, int pmstatus) {
    RCHK(need_particles(ctx, box));
    double f;
    RCHK(pm_factor(ctx, ef, pmstatus, &f));
    hipEvent_t* ev = box->ev;
    HIPCHK(ctx, hipEventRecord(ev[0], box->st));
    if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
    HIPCHK(ctx, hipEventRecord(ev[1], box->st));
    HIPCHK(ctx, hipEventRecord(ev[12], box->st));
    RCHK(lagpm_pm_deposit(ctx, box));
    HIPCHK(ctx, hipEventRecord(ev[2], box->st));
    RCHK(lagpm_pm_solve(ctx, box, ev + 3));
    RCHK(lagpm_pm_gradient(ctx, box));
    HIPCHK(ctx, hipEventRecord(ev[6], box->st));
    hipLaunchKernelGGL(set_factors_kernel, dim3(1), dim3(64), 0, box->st, box->dfac, 0.0, f, 0.0, 2);
    RCHK(lagpm_pm_interp(ctx, box));
    HIPCHK(ctx, hipEventRecord(ev[7], box->st));
    box->pm_timed = true;
    return LAGPM_OK;
}

}  // extern "C"

namespace {
// CorrMeasure (Cosmos/Correlation.c:13-116) of the box's particles: deposit, r2c and binning
// (allocating or from the device's static scratch), then normalisation, sigma_8 and xi(r) on the
// host as Cosmos/Correlation.c:80-103.
int power_spectrum_impl(lagpm_ctx* ctx, lagpm_box* box, double boxsize, double* pk, double* npk, double* cor,
                        double* sigma8, double* slope, bool noalloc) {
    RCHK(need_particles(ctx, box));
    if (!pk || !npk || !cor || !sigma8 || !slope || !(boxsize > 0))
        return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL output or non-positive box size");
    const int n = box->p.n;
    if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
    RCHK(lagpm_pm_deposit(ctx, box));
    RCHK(noalloc ? lagpm_pm_power_noalloc(ctx, box, pk, npk) : lagpm_pm_power(ctx, box, pk, npk));
    const double rth = 8.0 / boxsize * n, fact = 2 * M_PI * rth / (double)n;
    double var = 0;
    for (int i = 1; i <= n; i++) {
        pk[i] = npk[i] > 0 ? pk[i] / npk[i] : 0.0;
        const double x = fact * i, u = std::sin(x) - x * std::cos(x);
        var += pk[i] * (double)i * i * u * u / std::pow(x, 6);
    }
    pk[0] = npk[0] > 0 ? pk[0] / npk[0] : 0.0;
    *sigma8 = std::sqrt(4 * M_PI * var * 9.0);
    cor[0] = 0;
    for (int i = 1; i <= n; i++) {
        const double sep = 2 * M_PI * i / (double)n;
        double c = 0;
        for (int k = 1; k <= n; k++) c += std::sin(k * sep) * k * pk[k];
        cor[i] = 2.0 * n * c / i;
    }
    *slope = (n >= 3 && cor[3] > 0 && cor[2] > 0) ? (std::log(cor[3]) - std::log(cor[2])) / (std::log(4.0) - std::log(3.0)) : 0.0;
    return LAGPM_OK;
}
}  // namespace

extern "C" {

int lagpm_power_spectrum(lagpm_ctx* ctx, lagpm_box* box, double boxsize, double* pk, double* npk, double* cor,
                         double* sigma8, double* slope) {
    return power_spectrum_impl(ctx, box, boxsize, pk, npk, cor, sigma8, slope, false);
}

int lagpm_power_spectrum_noalloc(lagpm_ctx* ctx, lagpm_box* box, double boxsize, double* pk, double* npk, double* cor,
                                 double* sigma8, double* slope) {
    return power_spectrum_impl(ctx, box, boxsize, pk, npk, cor, sigma8, slope, true);
}

int lagpm_force_spline(const lagpm_params* p, double* diff, double* slope) {
    if (!p || !diff || !slope) return LAGPM_E_INVAL;
    if (check_params(nullptr, p) != LAGPM_OK) return LAGPM_E_INVAL;
    lagpm_build_tables(p, diff, slope);
    return LAGPM_OK;
}

// Queues sort (when needed), build and walk of one box. The FP32 walk runs on the context
// stream (its stacks are shared), so it waits for the box stream and the box stream for it.
// The FP32 walk writes its kick to the box's scratch arrays; with combine it is added to the
// velocities here, otherwise the caller adds it (the drift of a step) or drops it (autotuning).
static int tree_enqueue(lagpm_ctx* ctx, lagpm_box* box, double fact, bool combine) {
    hipEvent_t* ev = box->ev + 8;
    HIPCHK(ctx, hipEventRecord(ev[0], box->st));
    if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
    HIPCHK(ctx, hipEventRecord(ev[1], box->st));
    RCHK(lagpm_tree_build(ctx, box));
    HIPCHK(ctx, hipEventRecord(ev[2], box->st));
    if (box->p.precision == LAGPM_PREC_FP64) {
        HIPCHK(ctx, hipEventRecord(box->ev[14], box->st));
        RCHK(lagpm_tree_walk_f64(ctx, box, fact));
        HIPCHK(ctx, hipEventRecord(ev[3], box->st));
    } else {
        if (box->st != ctx->stream) HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ev[2], 0));
        RCHK(lagpm_tree_walk_f32(ctx, &box, &fact, 1));  // records ev[14] before the walk launch
        HIPCHK(ctx, hipEventRecord(ev[3], ctx->stream));
        if (box->st != ctx->stream) HIPCHK(ctx, hipStreamWaitEvent(box->st, ev[3], 0));
        if (combine) {
            hipLaunchKernelGGL(add3_kernel, dim3(nblocks(box->np, 256)), dim3(256), 0, box->st, box->vx, box->vy, box->vz,
                               (const double*)box->alt[0], (const double*)box->alt[1], (const double*)box->alt[2], box->np);
            HIPCHK(ctx, hipGetLastError());
        }
    }
    box->tree_timed = true;
    return LAGPM_OK;
}

static void tree_times(lagpm_box* box) {
    hipEvent_t* ev = box->ev + 8;
    float t01 = 0, t12 = 0, t23 = 0;
    (void)hipEventSynchronize(ev[3]);
    (void)hipEventElapsedTime(&t01, ev[0], ev[1]);
    (void)hipEventElapsedTime(&t12, ev[1], ev[2]);
    (void)hipEventElapsedTime(&t23, box->ev[14], ev[3]);  // walk start .. walk end
    box->stats.ms_build = t01 + t12;
    box->stats.ms_walk = t23;
}

static int tree_check(lagpm_ctx* ctx, lagpm_box* box) {
    if (!box->tab) return lagpm_fail(ctx, LAGPM_E_STATE, "box has no force tables");
    if (box->level_cell < 2)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "box too small for the cutoff: need n >= 4 * rsphere");
    return LAGPM_OK;
}

int lagpm_treecorrection(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact* ef) {
    RCHK(need_particles(ctx, box));
    if (!ef) return lagpm_fail(ctx, LAGPM_E_INVAL, "evolfact is NULL");
    RCHK(tree_check(ctx, box));
    if (box->np == 0) return LAGPM_OK;
    for (int attempt = 0;; attempt++) {
        RCHK(tree_enqueue(ctx, box, (double)ef->fact2, true));
        bool retry = false;
        int rc = lagpm_tree_finish(ctx, box, &retry);
        tree_times(box);
        if (rc == LAGPM_OK) return LAGPM_OK;
        // a build that overflowed its node arrays or met its depth limit applied no kick: grow
        // or deepen and redo it (a depth redo may still need larger node arrays)
        if (!retry || attempt > 1) return rc == LAGPM_E_OVERFLOW ? lagpm_fail(ctx, rc, "tree node arrays overflowed") : rc;
    }
}

int lagpm_tree_accel_sample(lagpm_ctx* ctx, lagpm_box* box, const int64_t* ids, size_t count, double* acc) {
    RCHK(need_particles(ctx, box));
    RCHK(tree_check(ctx, box));
    if (count == 0) return LAGPM_OK;
    if (!ids || !acc) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL ids or acc");
    if (count > box->np) return lagpm_fail(ctx, LAGPM_E_INVAL, "more sampled particles than the box holds");
    {
        std::vector<uint8_t> seen(box->np, 0);
        for (size_t t = 0; t < count; t++) {
            if (ids[t] < 0 || (size_t)ids[t] >= box->np)
                return lagpm_fail(ctx, LAGPM_E_INVAL, "sampled index " + std::to_string(ids[t]) + " out of range");
            if (seen[(size_t)ids[t]]++) return lagpm_fail(ctx, LAGPM_E_INVAL, "sampled index " + std::to_string(ids[t]) + " repeated");
        }
    }
    // the box's tree, built as the tree correction builds it (grown or deepened as needed)
    for (int attempt = 0;; attempt++) {
        HIPCHK(ctx, hipEventRecord(box->ev[8], box->st));
        if (!box->sorted) RCHK(lagpm_sort_particles(ctx, box));
        RCHK(lagpm_tree_build(ctx, box));
        bool retry = false;
        const int rc = lagpm_tree_finish(ctx, box, &retry);
        if (rc == LAGPM_OK) break;
        if (!retry || attempt > 1) return rc == LAGPM_E_OVERFLOW ? lagpm_fail(ctx, rc, "tree node arrays overflowed") : rc;
    }
    // scratch of this call (a diagnostic entry point: it allocates)
    int64_t* d_ids = nullptr;
    int* slot = nullptr;
    int2* groups = nullptr;
    uint32_t* sdsz = nullptr;
    double* d_acc = nullptr;
    int rc = lagpm_dalloc_bytes(ctx, (void**)&d_ids, count * sizeof(int64_t));
    if (rc == LAGPM_OK) rc = lagpm_dalloc_bytes(ctx, (void**)&slot, box->np * sizeof(int));
    if (rc == LAGPM_OK) rc = lagpm_dalloc_bytes(ctx, (void**)&groups, count * sizeof(int2));
    if (rc == LAGPM_OK) rc = lagpm_dalloc_bytes(ctx, (void**)&sdsz, DSZ_COUNT * sizeof(uint32_t));
    if (rc == LAGPM_OK) rc = lagpm_dalloc_bytes(ctx, (void**)&d_acc, 3 * count * sizeof(double));
    if (rc == LAGPM_OK && hipMemcpyAsync(d_ids, ids, count * sizeof(int64_t), hipMemcpyHostToDevice, box->st) != hipSuccess)
        rc = lagpm_fail(ctx, LAGPM_E_HIP, "upload of the sampled indices");
    if (rc == LAGPM_OK) rc = lagpm_tree_walk_sample(ctx, box, d_ids, count, slot, groups, sdsz, d_acc);
    unsigned long long c[4] = {0, 0, 0, 0};
    if (rc == LAGPM_OK) rc = lagpm_d2h(ctx, acc, d_acc, 3 * count * sizeof(double), box->st);
    if (rc == LAGPM_OK) rc = lagpm_d2h(ctx, c, box->counters, sizeof(c), box->st);
    (void)hipStreamSynchronize(box->st);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_ids); (void)hipFree(slot); (void)hipFree(groups); (void)hipFree(sdsz); (void)hipFree(d_acc);
    if (rc == LAGPM_OK && c[2]) rc = lagpm_fail(ctx, LAGPM_E_OVERFLOW, "tree walk stack overflow in " + std::to_string(c[2]) + " sampled particles");
    return rc;
}

int lagpm_box_set_tuning(lagpm_box* box, int nleaf, int group) {
    if (!box || nleaf < 4 || nleaf > 256 || group < 8 || group > 128) return LAGPM_E_INVAL;
    if (box->tune.nleaf != nleaf || box->tune.group != group) {
        lagpm_box_graphs_release(box);
        box->depth = LAGPM_KEY_BITS;  // another leaf size: find the tree depth again
    }
    box->tune.nleaf = nleaf;
    box->tune.group = group;
    return LAGPM_OK;
}

int lagpm_box_get_tuning(const lagpm_box* box, int* nleaf, int* group) {
    if (!box || !nleaf || !group) return LAGPM_E_INVAL;
    *nleaf = box->tune.nleaf;
    *group = box->tune.group;
    return LAGPM_OK;
}

int lagpm_box_set_group_mode(lagpm_box* box, int mode) {
    if (!box || (mode != LAGPM_GROUPS_TREE && mode != LAGPM_GROUPS_RUNS && mode != LAGPM_GROUPS_PACKED))
        return LAGPM_E_INVAL;
    const int runs = mode == LAGPM_GROUPS_RUNS ? 1 : (mode == LAGPM_GROUPS_PACKED ? 2 : 0);
    if (box->tune.runs != runs) lagpm_box_graphs_release(box);  // captured steps hold the mode
    box->tune.runs = runs;
    return LAGPM_OK;
}

int lagpm_box_get_group_mode(const lagpm_box* box) {
    if (!box) return LAGPM_E_INVAL;
    return box->tune.runs == 1 ? LAGPM_GROUPS_RUNS : (box->tune.runs == 2 ? LAGPM_GROUPS_PACKED : LAGPM_GROUPS_TREE);
}

int lagpm_box_autotune(lagpm_ctx* ctx, lagpm_box* box, int* nleaf, int* group) {
    RCHK(need_particles(ctx, box));
    RCHK(tree_check(ctx, box));
    if (box->np == 0) return LAGPM_OK;
    // (nleaf, group, group mode)
    static const int cand[][3] = {{16, 64, 0}, {32, 64, 0}, {8, 64, 0}, {16, 32, 0}, {32, 32, 0}, {8, 32, 0},
                                  {16, 16, 0}, {32, 16, 0}, {8, 16, 0}, {16, 96, 0}, {32, 96, 0},
                                  {16, 64, 1}, {32, 64, 1}, {16, 32, 1}, {32, 32, 1}, {16, 16, 1},
                                  {16, 64, 2}, {32, 64, 2}, {32, 96, 2}};
    double best = 1e300;
    int bl = box->tune.nleaf, bg = box->tune.group, bm = lagpm_box_get_group_mode(box);
    for (const auto& c : cand) {
        RCHK(lagpm_box_set_tuning(box, c[0], c[1]));
        RCHK(lagpm_box_set_group_mode(box, c[2]));
        double t = 1e300;
        for (int rep = 0; rep < 2; rep++) {  // the first run of a candidate may grow the node arrays
            RCHK(tree_enqueue(ctx, box, 0.0, false));
            bool retry = false;
            int rc = lagpm_tree_finish(ctx, box, &retry);
            if (rc != LAGPM_OK && !retry) return rc;
            if (rc == LAGPM_OK) {
                tree_times(box);
                t = std::min(t, box->stats.ms_build + box->stats.ms_walk);
            }
        }
        if (t < best) {
            best = t;
            bl = c[0];
            bg = c[1];
            bm = c[2];
        }
    }
    RCHK(lagpm_box_set_tuning(box, bl, bg));
    RCHK(lagpm_box_set_group_mode(box, bm));
    if (nleaf) *nleaf = bl;
    if (group) *group = bg;
    return LAGPM_OK;
}

int lagpm_tree_stats_get(const lagpm_box* box, lagpm_tree_stats* out) {
    if (!box || !out) return LAGPM_E_INVAL;
    *out = box->stats;
    out->depth_limit = box->depth;
    return LAGPM_OK;
}

int lagpm_box_timings(lagpm_ctx* ctx, lagpm_box* box, lagpm_timings* out) {
    if (!ctx || !box || !out) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    auto el = [&](int a, int b) {
        float t = 0;
        (void)hipEventSynchronize(box->ev[b]);
        (void)hipEventElapsedTime(&t, box->ev[a], box->ev[b]);
        return (double)t;
    };
    if (box->pm_timed) {
        out->ms_sort = el(0, 1);
        out->ms_tsc = el(12, 2);
        out->ms_fft_r2c = el(2, 3);
        out->ms_green = el(3, 4);
        out->ms_fft_c2r = el(4, 5);
        out->ms_grad = el(5, 6);
        out->ms_interp = el(6, 7);
    }
    if (box->tree_timed) {
        out->ms_tree_sort = el(8, 9);
        out->ms_build = el(9, 10);
        out->ms_walk = el(14, 11);  // from right before the walk launch to its end
    }
    return LAGPM_OK;
}

int lagpm_evolfactor(const lagpm_params* p, double anow, double astep, lagpm_evolfact* out) {
    if (!p || !out || !(anow > 0) || !(p->omep > 0) || !(p->amax > 0)) return LAGPM_E_INVAL;
    if (!(anow - 0.5 * astep > 0)) return LAGPM_E_INVAL;
    cosmo c{p->omep, p->omeplam, p->amax};
    const int n = p->n;
    out->afact = out->bfact = 0;
    out->fact1 = (float)fact1_of(c, anow - 0.5 * astep, anow + 0.5 * astep);
    out->fact2 = (float)fact2_of(c, n, anow - 0.5 * astep, anow + 0.5 * astep);
    out->pfact = (float)(astep * n);
    out->fact1_push = (float)fact1_of(c, anow, anow + 0.5 * astep);
    out->fact1_pull = (float)fact1_of(c, anow - 0.5 * astep, anow);
    out->fact2_push = (float)fact2_of(c, n, anow, anow + 0.5 * astep);
    out->fact2_pull = (float)fact2_of(c, n, anow - 0.5 * astep, anow);
    return LAGPM_OK;
}

int lagpm_hubble_drag(lagpm_ctx* ctx, lagpm_box* box, double fact1) {
    RCHK(need_particles(ctx, box));
    if (box->np == 0) return LAGPM_OK;
    hipLaunchKernelGGL(set_factors_kernel, dim3(1), dim3(64), 0, box->st, box->dfac, fact1, 0.0, 0.0, 1);
    hipLaunchKernelGGL(scale3_kernel, dim3(nblocks(box->np, 256)), dim3(256), 0, box->st, box->vx, box->vy, box->vz,
                       box->np, (const double*)box->dfac);
    HIPCHK(ctx, hipGetLastError());
    return LAGPM_OK;
}

int lagpm_drift(lagpm_ctx* ctx, lagpm_box* box, double pfact) {
    RCHK(need_particles(ctx, box));
    if (box->np == 0) return LAGPM_OK;
    hipLaunchKernelGGL(set_factors_kernel, dim3(1), dim3(64), 0, box->st, box->dfac, 0.0, 0.0, pfact, 4);
    hipLaunchKernelGGL(drift_kernel, dim3(nblocks(box->np, 256)), dim3(256), 0, box->st, box->x, box->y, box->z, box->vx,
                       box->vy, box->vz, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr, box->np,
                       (const double*)box->dfac, (double)box->p.n, (const uint32_t*)nullptr);
    HIPCHK(ctx, hipGetLastError());
    box->sorted = false;
    return LAGPM_OK;
}

}  // extern "C"

namespace {

#ifndef LAGPM_BATCH_GROWTH
#define LAGPM_BATCH_GROWTH 1.5
#endif

int ensure_pool(lagpm_ctx* ctx, size_t want) {
    while (ctx->pool.size() < want) {
        hipStream_t s = nullptr;
        HIPCHK(ctx, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ctx->pool.push_back(s);
    }
    return LAGPM_OK;
}

// One box's step before the walk, queued on box->st with the factors already in box->dfac:
// Hubble drag and sort, then a fork: the PM kick (deposit, FFTs, Green, gradient, interpolation)
// on box->st beside the tree build on box->st2, joined on box->st. The PM branch is queued first:
// a graph launch and an eager run both submit in queue order, and the build's many small kernels
// would otherwise hold back the start of the longer PM chain. The build's latency-bound kernels
// and the PM's memory-bound ones share the GPU. With ev: phase events (eager runs only; a capture
// records none, but the fork and join events are part of it).
int front_body(lagpm_ctx* ctx, lagpm_box* box, hipEvent_t* ev) {
    hipStream_t st = box->st, st2 = box->st2;
    if (ev) HIPCHK(ctx, hipEventRecord(ev[0], st));
    // the Hubble drag (v *= dfac[0]) rides in the sort's gather; a lone box gathers its
    // velocities on st3, beside the deposit and the build, and the PM kick waits for them
    const bool vsplit = box->st3 && box->st3 != st;
    RCHK(lagpm_sort_particles(ctx, box, (const double*)box->dfac, vsplit ? box->st3 : nullptr));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[1], st));
    HIPCHK(ctx, hipEventRecord(box->evfork, st));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[12], st));
    RCHK(lagpm_pm_deposit(ctx, box));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[2], st));
    RCHK(lagpm_pm_solve(ctx, box, ev ? ev + 3 : nullptr));
    RCHK(lagpm_pm_gradient(ctx, box));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[6], st));
    if (vsplit) HIPCHK(ctx, hipStreamWaitEvent(st, box->evvel, 0));
    RCHK(lagpm_pm_interp(ctx, box));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[7], st));
    const bool fork = st2 != st;  // several boxes in a batch build on their own streams
    if (fork) HIPCHK(ctx, hipStreamWaitEvent(st2, box->evfork, 0));
    if (ev) {
        HIPCHK(ctx, hipEventRecord(ev[8], st2));
        HIPCHK(ctx, hipEventRecord(ev[9], st2));  // sorted above: no tree sort
    }
    box->st = st2;
    const int rb = lagpm_tree_build(ctx, box);
    box->st = st;
    RCHK(rb);
    HIPCHK(ctx, hipEventRecord(box->ev[10], st2));  // build done: the join
    if (fork) HIPCHK(ctx, hipStreamWaitEvent(st, box->ev[10], 0));
    return LAGPM_OK;
}

struct box_ptrs {  // host-side pointer state a front changes (sort swaps)
    double *x, *y, *z, *vx, *vy, *vz, *alt[6];
    uint32_t *id, *id_alt;
    int parity;
    bool sorted;
};
box_ptrs save_ptrs(const lagpm_box* b) {
    box_ptrs s{b->x, b->y, b->z, b->vx, b->vy, b->vz, {}, b->id, b->id_alt, b->parity, b->sorted};
    for (int a = 0; a < 6; a++) s.alt[a] = b->alt[a];
    return s;
}
void load_ptrs(lagpm_box* b, const box_ptrs& s) {
    b->x = s.x; b->y = s.y; b->z = s.z; b->vx = s.vx; b->vy = s.vy; b->vz = s.vz;
    for (int a = 0; a < 6; a++) b->alt[a] = s.alt[a];
    b->id = s.id; b->id_alt = s.id_alt; b->parity = s.parity; b->sorted = s.sorted;
}
// what a front does to the host-side state, without queueing anything (graph replays)
void front_host_effects(lagpm_box* b) {
    std::swap(b->x, b->alt[0]); std::swap(b->y, b->alt[1]); std::swap(b->z, b->alt[2]);
    std::swap(b->vx, b->alt[3]); std::swap(b->vy, b->alt[4]); std::swap(b->vz, b->alt[5]);
    std::swap(b->id, b->id_alt);
    b->parity ^= 1;
    b->sorted = true;
}

// Everything of one box's step before the walk, queued on box->st (and box->st2), then the event
// the walk waits for (ev[13]). With graphs on, the front is captured once per particle-set parity
// and replayed; factors are written first.
int step_front(lagpm_ctx* ctx, lagpm_box* box, const lagpm_evolfact& ef) {
    hipLaunchKernelGGL(set_factors_kernel, dim3(1), dim3(64), 0, box->st, box->dfac, (double)ef.fact1, (double)ef.fact2,
                       (double)ef.pfact, 7);
    HIPCHK(ctx, hipGetLastError());
    auto eager = [&]() -> int {
        RCHK(front_body(ctx, box, box->ev));
        box->pm_timed = true;
        HIPCHK(ctx, hipEventRecord(box->ev[13], box->st));
        return LAGPM_OK;
    };
    if (!ctx->use_graphs || box->graph_failed) return eager();
    // everything a capture must not do (plan creation and its check synchronise) happens first
    RCHK(lagpm_pm_prepare(ctx, box));
    HIPCHK(ctx, hipEventRecord(box->ev[8], box->st));
    HIPCHK(ctx, hipEventRecord(box->ev[9], box->st));
    hipGraphExec_t& ex = box->gfront[box->parity];
    if (!ex) {
        const box_ptrs before = save_ptrs(box);
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(box->st, hipStreamCaptureModeRelaxed) == hipSuccess;
        int rc = ok ? front_body(ctx, box, nullptr) : LAGPM_E_HIP;
        if (ok) ok = hipStreamEndCapture(box->st, &g) == hipSuccess && rc == LAGPM_OK && g;
        if (ok) ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        load_ptrs(box, before);
        (void)hipGetLastError();
        if (!ok) {  // run eagerly from now on for this box
            ex = nullptr;
            box->graph_failed = true;
            return eager();
        }
    }
    HIPCHK(ctx, hipGraphLaunch(ex, box->st));
    front_host_effects(box);
    box->pm_timed = false;  // no phase events inside a graph
    HIPCHK(ctx, hipEventRecord(box->ev[10], box->st));
    HIPCHK(ctx, hipEventRecord(box->ev[13], box->st));
    return LAGPM_OK;
}

// The drift of a step; an FP32 box's tree kick (written by the walk to the scratch arrays) is
// added to the velocities first.
int drift_enqueue(lagpm_ctx* ctx, lagpm_box* box, double pfact, bool guarded) {
    if (!guarded)  // a redone step: the factor of this step is already in dfac[2] otherwise
        hipLaunchKernelGGL(set_factors_kernel, dim3(1), dim3(64), 0, box->st, box->dfac, 0.0, 0.0, pfact, 4);
    const bool tk = box->p.precision == LAGPM_PREC_FP32;
    hipLaunchKernelGGL(drift_kernel, dim3(nblocks(box->np, 256)), dim3(256), 0, box->st, box->x, box->y, box->z, box->vx,
                       box->vy, box->vz, tk ? (const double*)box->alt[0] : nullptr,
                       tk ? (const double*)box->alt[1] : nullptr, tk ? (const double*)box->alt[2] : nullptr, box->np,
                       (const double*)box->dfac, (double)box->p.n,
                       guarded ? (const uint32_t*)box->dsz : (const uint32_t*)nullptr);
    HIPCHK(ctx, hipGetLastError());
    box->sorted = false;
    return LAGPM_OK;
}

// Queues step_front for each (box, factors) pair (the boxes already carry their streams).
int queue_fronts(lagpm_ctx* ctx, const std::vector<lagpm_box*>& bs, const std::vector<const lagpm_evolfact*>& efs) {
    for (size_t i = 0; i < bs.size(); i++) RCHK(step_front(ctx, bs[i], *efs[i]));
    return LAGPM_OK;
}

int batch_step(lagpm_ctx* ctx, lagpm_box* const* boxes, int nbox, const double* anow, const double* astep) {
    if (!ctx || (nbox > 0 && (!boxes || !anow || !astep)) || nbox < 0)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL argument or negative box count");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<lagpm_evolfact> ef((size_t)nbox);
    std::vector<lagpm_box*> live;
    for (int i = 0; i < nbox; i++) {
        lagpm_box* b = boxes[i];
        if (!b) return lagpm_fail(ctx, LAGPM_E_INVAL, "NULL box");
        if (!b->uploaded) return lagpm_fail(ctx, LAGPM_E_STATE, "no particles uploaded");
        RCHK(tree_check(ctx, b));
        for (int j = 0; j < i; j++)
            if (boxes[j] == b) return lagpm_fail(ctx, LAGPM_E_INVAL, "a box appears twice in one batch");
        if (lagpm_evolfactor(&b->p, anow[i], astep[i], &ef[(size_t)i]) != LAGPM_OK)
            return lagpm_fail(ctx, LAGPM_E_INVAL, "bad scale factor or step for box " + std::to_string(i));
        if (b->np) live.push_back(b);
    }
    if (live.empty()) return LAGPM_OK;
    const size_t S = std::min<size_t>(live.size(), 8);
    RCHK(ensure_pool(ctx, std::max<size_t>(S, 2)));  // a lone box uses a second pool stream
    // the pool streams start after everything queued on the context stream so far
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (size_t k = 0; k < S; k++) HIPCHK(ctx, hipStreamWaitEvent(ctx->pool[k], ctx->ev[0], 0));
    std::vector<double> fact((size_t)nbox);
    for (int i = 0; i < nbox; i++) fact[(size_t)i] = ef[(size_t)i].fact2;
    // Pipelined batch: all boxes FP32 with one force table, enough of them for the stages. The
    // boxes are cut into `stages` consecutive groups of growing particle shares; group g's walk
    // (one fused launch, leaving ctx->stage_reserve_cu CUs free) runs while the fronts of group
    // g+1 are on the pool streams, whose small kernels are latency bound and fill those CUs. Every
    // drift is queued after all fronts, so no front waits behind a walk on its pool stream.
    int stages = std::max(1, ctx->batch_stages);
    for (lagpm_box* b : live)
        if (b->p.precision != LAGPM_PREC_FP32 || b->tab != live[0]->tab) stages = 1;
    stages = std::min<int>(stages, (int)(live.size() / 2));  // at least two boxes per group
    if (stages < 2) stages = 1;
    int rc = LAGPM_OK;
    if (stages > 1) {
        size_t ptot = 0;
        for (lagpm_box* b : live) ptot += b->np;
        size_t k = 0, acc = 0, li = 0;
        int g = 0;
        std::vector<lagpm_box*> grp;
        std::vector<double> gf;
        std::vector<const lagpm_evolfact*> gef;
        for (int i = 0; i < nbox && rc == LAGPM_OK; i++) {
            lagpm_box* b = boxes[i];
            if (!b->np) continue;
            b->st = ctx->pool[k++ % S];
            b->st2 = b->st;
            b->st3 = b->st;
            grp.push_back(b);
            gf.push_back(fact[(size_t)i]);
            gef.push_back(&ef[(size_t)i]);
            acc += b->np;
            li++;
            const bool last = li == live.size();
            // close the group at its share of the particles (the last group takes the rest):
            // shares grow by LAGPM_BATCH_GROWTH per group, so the first group's front, which
            // nothing overlaps, is short, and each later group's front fits beside the previous
            // group's (longer) walk
            const double gr = LAGPM_BATCH_GROWTH;
            const double upto = gr == 1.0 ? (double)(g + 1) / stages : (std::pow(gr, g + 1) - 1.0) / (std::pow(gr, stages) - 1.0);
            if (last || (g < stages - 1 && (double)acc >= upto * (double)ptot)) {
                rc = queue_fronts(ctx, grp, gef);
                for (lagpm_box* c : grp)
                    if (hipStreamWaitEvent(ctx->stream, c->ev[13], 0) != hipSuccess) rc = LAGPM_E_HIP;
                if (rc == LAGPM_OK)
                    rc = lagpm_tree_walk_f32(ctx, grp.data(), gf.data(), (int)grp.size(), last ? 0 : ctx->stage_reserve_cu);
                for (lagpm_box* c : grp)
                    if (rc == LAGPM_OK && hipEventRecord(c->ev[11], ctx->stream) != hipSuccess) rc = LAGPM_E_HIP;
                grp.clear();
                gf.clear();
                gef.clear();
                g++;
            }
        }
        for (int i = 0; i < nbox && rc == LAGPM_OK; i++) {
            lagpm_box* b = boxes[i];
            if (!b->np) continue;
            if (hipStreamWaitEvent(b->st, b->ev[11], 0) != hipSuccess) rc = LAGPM_E_HIP;
            b->tree_timed = true;
            if (rc == LAGPM_OK) rc = drift_enqueue(ctx, b, (double)ef[(size_t)i].pfact, true);
        }
    } else {
        size_t k = 0;
        std::vector<lagpm_box*> fb;
        std::vector<const lagpm_evolfact*> fe;
        for (int i = 0; i < nbox; i++) {
            lagpm_box* b = boxes[i];
            if (!b->np) continue;
            // A lone box forks its build onto the context stream (a second pool stream can share the
            // box stream's hardware queue, 4 per process, which would serialise the two branches).
            // Several boxes already run side by side on the pool streams: each keeps its build on its
            // own stream, as one shared build stream would serialise the builds of all boxes.
            b->st = ctx->pool[k++ % S];
            b->st2 = live.size() == 1 ? ctx->stream : b->st;
            b->st3 = live.size() == 1 ? ctx->pool[1] : b->st;  // a lone box's velocity gather
            fb.push_back(b);
            fe.push_back(&ef[(size_t)i]);
        }
        rc = queue_fronts(ctx, fb, fe);
        // FP32 walks: one persistent launch per force table on the context stream
        std::vector<int> done((size_t)nbox, 0);
        for (int i = 0; i < nbox && rc == LAGPM_OK; i++) {
            lagpm_box* b = boxes[i];
            if (!b->np || done[(size_t)i]) continue;
            if (b->p.precision == LAGPM_PREC_FP64) {
                if (hipEventRecord(b->ev[14], b->st) != hipSuccess) rc = LAGPM_E_HIP;
                if (rc == LAGPM_OK) rc = lagpm_tree_walk_f64(ctx, b, fact[(size_t)i]);
                if (rc == LAGPM_OK && hipEventRecord(b->ev[11], b->st) != hipSuccess) rc = LAGPM_E_HIP;
                done[(size_t)i] = 1;
                continue;
            }
            std::vector<lagpm_box*> cls;
            std::vector<double> cf;
            for (int j = i; j < nbox; j++) {
                lagpm_box* c = boxes[j];
                if (!c->np || done[(size_t)j] || c->p.precision != LAGPM_PREC_FP32 || c->tab != b->tab) continue;
                if (hipStreamWaitEvent(ctx->stream, c->ev[13], 0) != hipSuccess) rc = LAGPM_E_HIP;
                cls.push_back(c);
                cf.push_back(fact[(size_t)j]);
                done[(size_t)j] = 1;
            }
            if (rc == LAGPM_OK) rc = lagpm_tree_walk_f32(ctx, cls.data(), cf.data(), (int)cls.size());
            if (rc == LAGPM_OK && hipEventRecord(ctx->ev[1], ctx->stream) != hipSuccess) rc = LAGPM_E_HIP;
            for (lagpm_box* c : cls) {
                if (rc != LAGPM_OK) break;
                if (hipEventRecord(c->ev[11], ctx->stream) != hipSuccess || hipStreamWaitEvent(c->st, ctx->ev[1], 0) != hipSuccess)
                    rc = LAGPM_E_HIP;
            }
        }
        for (int i = 0; i < nbox && rc == LAGPM_OK; i++) {
            lagpm_box* b = boxes[i];
            if (!b->np) continue;
            b->tree_timed = true;
            rc = drift_enqueue(ctx, b, (double)ef[(size_t)i].pfact, true);
        }
    }
    // completion: sizes and counters of every box; a box whose node arrays overflowed got
    // neither tree kick nor drift, so it is rebuilt with larger arrays and finished here
    int first_err = rc;
    // every box's sizes and counters in one wait: the copies follow each box's work on its stream
    tree_readback* rb = nullptr;
    if (first_err == LAGPM_OK) {
        if (ctx->readback_count < (size_t)nbox) {
            if (ctx->readback) (void)hipHostFree(ctx->readback);
            ctx->readback = nullptr;
            ctx->readback_count = 0;
            const hipError_t e = hipHostMalloc(&ctx->readback, sizeof(tree_readback) * (size_t)nbox, hipHostMallocDefault);
            if (e != hipSuccess) first_err = lagpm_fail(ctx, LAGPM_E_HIP, std::string("pinned readback: ") + hipGetErrorString(e));
            else ctx->readback_count = (size_t)nbox;
        }
        rb = static_cast<tree_readback*>(ctx->readback);
        for (int i = 0; i < nbox && first_err == LAGPM_OK; i++)
            if (boxes[i]->np) first_err = lagpm_tree_readback_enqueue(ctx, boxes[i], rb + i);
        for (size_t q = 0; q < ctx->pool.size() && first_err == LAGPM_OK; q++)
            if (hipStreamSynchronize(ctx->pool[q]) != hipSuccess) first_err = LAGPM_E_HIP;
        if (first_err == LAGPM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) first_err = LAGPM_E_HIP;
    }
    for (int i = 0; i < nbox; i++) {
        lagpm_box* b = boxes[i];
        if (!b->np) continue;
        if (first_err == LAGPM_OK) {
            bool retry = false;
            int r = lagpm_tree_finish_with(ctx, b, rb[i], &retry);
            // redo (at most twice: a depth redo may still need larger node arrays)
            for (int attempt = 0; retry && attempt < 2; attempt++) {
                retry = false;
                b->st = ctx->stream;
                HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
                r = tree_enqueue(ctx, b, fact[(size_t)i], false);
                if (r == LAGPM_OK) r = lagpm_tree_finish(ctx, b, &retry);
                if (r == LAGPM_OK) r = drift_enqueue(ctx, b, (double)ef[(size_t)i].pfact, false);
            }
            if (r != LAGPM_OK && retry) r = lagpm_fail(ctx, LAGPM_E_OVERFLOW, "tree node arrays overflowed");
            if (r == LAGPM_OK) tree_times(b);
            if (r != LAGPM_OK) first_err = r;
        }
        b->st = ctx->stream;
    }
    for (size_t q = 0; q < S; q++) (void)hipStreamSynchronize(ctx->pool[q]);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return first_err;
}

}  // namespace

void lagpm_box_graphs_release(lagpm_box* box) {
    for (int q = 0; q < 2; q++)
        if (box->gfront[q]) {
            (void)hipGraphExecDestroy(box->gfront[q]);
            box->gfront[q] = nullptr;
        }
}

extern "C" {

int lagpm_batch_step(lagpm_ctx* ctx, lagpm_box* const* boxes, int nbox, const double* anow, const double* astep) {
    return batch_step(ctx, boxes, nbox, anow, astep);
}

int lagpm_set_step_graphs(lagpm_ctx* ctx, int enable) {
    if (!ctx) return LAGPM_E_INVAL;
    ctx->use_graphs = enable != 0;
    return LAGPM_OK;
}

int lagpm_set_walk_census(lagpm_ctx* ctx, int enable) {
    if (!ctx) return LAGPM_E_INVAL;
    ctx->walk_census = enable != 0;
    return LAGPM_OK;
}

int lagpm_set_batch_pipeline(lagpm_ctx* ctx, int stages, int reserve_cu) {
    if (!ctx) return LAGPM_E_INVAL;
    if (stages < 1 || stages > 16 || reserve_cu < 0 || reserve_cu >= ctx->num_cu)
        return lagpm_fail(ctx, LAGPM_E_INVAL, "stages must be in [1, 16] and reserve_cu in [0, CUs)");
    ctx->batch_stages = stages;
    ctx->stage_reserve_cu = reserve_cu;
    return LAGPM_OK;
}

int lagpm_kdk_step(lagpm_ctx* ctx, lagpm_box* box, double anow, double astep) {
    return batch_step(ctx, &box, 1, &anow, &astep);
}

}  // extern "C"
