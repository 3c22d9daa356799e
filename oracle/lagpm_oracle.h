/*
 * lagpm_oracle.h — CPU oracle for the TreePM gravity step.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the timed
 * CPU baseline). The product path (lageunha_amd, liblagpm) never links or calls it.
 *
 * Mesh layout: idx = i + n*(j + n*k), x fastest. Spectrum layout (r2c): complex
 * float pairs at (i + (n/2+1)*(j + n*k)), i in [0, n/2].
 */
#ifndef LAGPM_ORACLE_H
#define LAGPM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#define ORC_NSPLINE 4096
#define ORC_QUIRK_D1 1  /* AddNode adds the x component three times, OST/ost.ver2.c:273-275 */
#define ORC_QUIRK_D2 2  /* quad[5] accumulates m*y*x, OST/ost.ver2.c:75 */
#define ORC_QUIRK_D3 4  /* DirectSummation adds into pos[0], Cosmos/treecorrection.c:209-217 */

int    orc_num_threads(void);
void   orc_set_threads(int nthreads);

/* TSC deposit, Cosmos/pmmain.h:3-78. */
void   orc_tsc_fixed(int n, size_t np, const double* x, const double* y, const double* z,
                     float mass, int scale_bits, int64_t* acc);
void   orc_tsc_float(int n, size_t np, const double* x, const double* y, const double* z,
                     float mass, float* den);
int    orc_fixed_scale_bits(double total_mass);
void   orc_fixed_to_density(int n, const int64_t* acc, int scale_bits, float* den);

/* 3-D FFT in double precision (own radix-2 / direct DFT). sign=-1 forward. */
void   orc_fft3d(int n, double* data /* complex interleaved, n^3 */, int sign);
void   orc_r2c(int n, const double* in, double* out /* n*n*(n/2+1) complex */);
void   orc_c2r(int n, const double* in, double* out);

/* Green multiply on a float spectrum, Cosmos/psolver.c:13-86. */
void   orc_psolver_f32(int n, double efold, int deconvolve, float* spec);
void   orc_psolver_f64(int n, double efold, int deconvolve, double* spec);

/* FDA4 gradient mesh (fx, fy, fz, phi) per cell, Cosmos/fda4.h:302-384, float. */
void   orc_fda4_gradient(int n, const float* phi, float* grad4);
/* FDA4 interpolation kick in the reference's float order, fda4.h:11-102,276-300. dv = fact2*F. */
void   orc_fda4_kick(int n, const float* phi, size_t np, const double* x, const double* y,
                     const double* z, float fact2, float* dvx, float* dvy, float* dvz);
void   orc_fda4_kick_f64(int n, const double* phi, size_t np, const double* x, const double* y,
                         const double* z, double fact2, double* dv /* 3*np */);

/* Full PM force chain. prec 0: reference storage (float mesh, float kick); prec 1: FP64. */
void   orc_pm_accel(int n, size_t np, const double* x, const double* y, const double* z,
                    double mass, double efold, int deconvolve, int prec, double* acc /*3*np*/);
/* the mesh of all particles, the kick interpolated at targets[0..ntarget) only (acc: 3*ntarget) */
void   orc_pm_accel_at(int n, size_t np, const double* x, const double* y, const double* z, double mass,
                       double efold, int deconvolve, int prec, size_t ntarget, const int64_t* targets, double* acc);
/* the same with the wall time of each phase: sec[0..4] = TSC, F-FFT, psolv, B-FFT, 4FDA */
void   orc_pm_accel_timed(int n, size_t np, const double* x, const double* y, const double* z,
                          double mass, double efold, int deconvolve, int prec, double* acc, double* sec);

/* Force tables, OST/Model4TreeForce.c:48-116 (law 0), erfc law (law 1), or
 * ForceLaw/treeforce.c:134-186 (law 2: eps fixed at 0.1, step 0.2 bin, entry 0 zero). */
void   orc_force_spline(int law, int n, double eps, double rs, double* diff, double* slope);
/* The same with the derivative step dr_bins (in bins) chosen by the caller. */
void   orc_force_spline_dr(int law, int n, double eps, double rs, double dr_bins, double* diff, double* slope);
double orc_pmcf(int n, double eps, double r);
double orc_erfc_fr(int n, double eps, double rs, double r);
float  orc_point_grav(const double* diff, const double* slope, float dist);
void   orc_cell_grav(const double* diff, const double* slope, float dist, float* f1, float* f2, float* f3);

/* Exact short-range pair sum, minimum image, cutoff rcut, reference table lookup. */
void   orc_pp_direct(int n, size_t np, const double* x, const double* y, const double* z,
                     float mass, const double* diff, const double* slope, double rcut,
                     size_t ntarget, const int64_t* targets, double* acc /* 3*ntarget */);
/* Same law evaluated analytically in FP64 (no table): fr(r)/r times r-vector. */
void   orc_pp_direct_analytic(int law, int n, double eps, double rs, size_t np,
                              const double* x, const double* y, const double* z, double mass,
                              double rcut, size_t ntarget, const int64_t* targets, double* acc);

/* Reference tree correction: linked cells of width cellw, per-cell octree over the
 * 27 neighbour cells, reference walk. dv = fact2 * a. quirks: ORC_QUIRK_* bitmask. */
int    orc_treecorrection_ref(int n, size_t np, const double* x, const double* y,
                              const double* z, float mass, const double* diff,
                              const double* slope, double rsphere, double theta,
                              double cellw, float fact2, int quirks, float* dv /* 3*np */);

/* The same kicks for the particles with sel[p] != 0 only (others get 0). */
int    orc_treecorrection_ref_sel(int n, size_t np, const double* x, const double* y,
                                  const double* z, float mass, const double* diff,
                                  const double* slope, double rsphere, double theta,
                                  double cellw, float fact2, int quirks, const uint8_t* sel,
                                  float* dv /* 3*np */);

/* Quadrupole node interaction of the reference (AddNode without D1), for pinning. */
void   orc_node_accel(const double* diff, const double* slope, double rcut,
                      size_t ns, const double* sx, const double* sy, const double* sz,
                      const float* sm, const double* target, int quirks, float* acc3);

/* Ewald sum of the periodic Newtonian force, G times masses, box side L. */
void   orc_ewald_accel(size_t np, const double* x, const double* y, const double* z,
                       const double* m, double L, double G, double alpha, int nreal, int kmax,
                       double* acc);
/* The same with the minimum-image pair force Plummer-softened (eps > 0): the periodic force of
 * the TreePM split's target law, r / (r^2 + eps^2)^{3/2} at short range. */
void   orc_ewald_accel_soft(size_t np, const double* x, const double* y, const double* z,
                            const double* m, double L, double G, double alpha, int nreal, int kmax,
                            double eps, double* acc);

/* Kick/drift factors, Cosmos/CosmosEvolFactor.c:230-256, FP64 quadrature. */
typedef struct orc_evolfact {
    double afact, bfact, pfact, fact1, fact2, fact1_push, fact2_push, fact1_pull, fact2_pull;
} orc_evolfact;
double orc_hofez(double omep, double omeplam, double amax, double a);
double orc_ap(double omep, double omeplam, double amax, double a);
double orc_fact1(double omep, double omeplam, double amax, double a1, double a2);
double orc_fact2(double omep, double omeplam, double amax, int n, double a1, double a2);
void   orc_evolfactor(double omep, double omeplam, double amax, int n, double anow,
                      double astep, orc_evolfact* out);

/* CorrMeasure (Cosmos/Correlation.c:13-116) on a density-contrast spectrum. */
void   orc_power_spectrum(int n, const double* spec, double boxsize, double* pk /*n+1*/, double* npk /*n+1*/,
                          double* cor /*n+1*/, double* sigma8, double* slope);

/* PreFoF halo-candidate flags (Cosmos/PreFoF.c:267-386): particles are binned into cells of
 * width 4 (Cosmos/BasicCell.h:1) over [0, n); friends-of-friends groups with linking length
 * fof_link are found inside each cell alone (OST/fof.c:220-293: a pair links when
 * sqrt(dx^2 + dy^2 + dz^2) <= fof_link in FP64, coordinates relative to the cell corner); a group
 * is kept (its particles flagged) unless it lies inside [fof_link, 4 - fof_link) on every axis
 * with fewer than minnum members (CountHaloCandidates, PreFoF.c:207-265). Returns the number of
 * flagged particles; flag[i] is 0 or 1. Brute-force pairs per cell. */
size_t orc_prefof(int n, size_t np, const double* x, const double* y, const double* z, float fof_link,
                  int minnum, uint8_t* flag);

#endif
