/*
 * g3l_oracle.h -- CPU oracle for the NNMap / NMapMap integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in g3lhalo_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never links, calls or falls back to it.
 *
 * It restates, in FP64 C++ with OpenMP over the points of each batch (the
 * reference CPU build's structure, src/NNMap_Model.cu:519), the reference's
 *   - h-adaptive cubature driver  (src/hcubature.cpp:115-1067, src/converged.h)
 *   - integrands of the 1/2/3-halo NNMap terms (src/NNMap_Model.cu:544-1108)
 *   - integrands of the 1/2/3-halo NMapMap terms (src/NMapMap_Model.cu:263-729)
 *   - HOD occupations and the shared primitives (src/HOD.cu:35-94,
 *     src/generalFunctions.cu:6-244)
 *   - model set-up: integral borders, rho_bar(z) and n_bar(z)
 *     (src/NNMap_Model.cu:15-101, 244-297; src/NMapMap_Model.cu:14-167)
 * with the reference's floating-point evaluation order kept expression by
 * expression, so that with the same libm the values are bit-identical.
 *
 * Parity pinning (which check pins which function) is listed in the header
 * comment of g3l_oracle.cpp.
 */
#ifndef G3L_ORACLE_H
#define G3L_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same signature as the reference's integrand_v (src/cubature.h:55-57). */
typedef int (*orc_integrand_v)(unsigned ndim, size_t npt, const double *x,
                               void *fdata, unsigned fdim, double *fval);

/* Same signature as the reference's hcubature_v (src/cubature.h:90-94);
 * the error_norm enum is passed as int. */
typedef int (*orc_hcubature_v_fn)(unsigned fdim, orc_integrand_v f, void *fdata,
                                  unsigned dim, const double *xmin, const double *xmax,
                                  size_t maxEval, double reqAbsError, double reqRelError,
                                  int norm, double *val, double *err);

/* Statistic selector and term numbers. */
enum { ORC_NNMAP = 0, ORC_NMAPMAP = 1 };

typedef struct orc_stats {
  uint64_t neval;      /* integrand evaluations (sum of npts over calls) */
  uint32_t ncalls;     /* integrand callback invocations (Gladwell batches) */
  uint32_t trace_len;  /* entries written to trace */
} orc_stats;

/* The oracle's own restatement of hcubature_v. */
int orc_hcubature_v(unsigned fdim, orc_integrand_v f, void *fdata, unsigned dim,
                    const double *xmin, const double *xmax, size_t maxEval,
                    double reqAbsError, double reqRelError, int norm,
                    double *val, double *err);

/* Model: tables are copied (P_lin gets nbins trailing zeros, see the
 * P(k) look-up note in g3l_oracle.cpp). */
void *orc_model_create(int nbins, double zmin, double zmax, double kmin, double kmax,
                       double mmin, double mmax, const double *g, const double *p_lens1,
                       const double *p_lens2, const double *w, const double *dwdz,
                       const double *hmf, const double *P_lin, const double *b_h,
                       const double *concentration, double H0, double Omega_m,
                       double Omega_Lambda);
void orc_model_destroy(void *model);

/* hod = {f, alpha, M_th, sigma, M', beta} (src/HOD.cu:15-20).
 * identical != 0 means both HOD slots are one object (src/HOD.cu:81-83).
 * hcub == NULL uses the oracle's own cubature. */
int orc_model_set_hods(void *model, const double *hod1, const double *hod2, double A,
                       double epsilon, int identical, orc_hcubature_v_fn hcub);
int orc_model_densities(void *model, double *rho_bar, double *n_bar1, double *n_bar2);
int orc_model_borders(void *model, double *out4); /* zmin_i, zmax_i, mmin_i, mmax_i */

/* One adaptive integral of one term.  out[0] = prefactored result (what
 * NNMap_1h etc. return), out[1] = raw integral, out[2] = raw error. */
int orc_eval_term(void *model, int stat, int term, double theta1, double theta2,
                  double theta3, int type1, int type2, double reltol, uint64_t maxeval,
                  orc_hcubature_v_fn hcub, double *out, orc_stats *stats,
                  uint32_t *trace, uint32_t trace_cap);

/* Raw integrand at npts points x[npts][ndim] (ndim = 4 + term; for the
 * n_bar integrand use term 0, ndim 1, theta1 = z and type1 = HOD slot). */
int orc_integrand(void *model, int stat, int term, double theta1, double theta2,
                  double theta3, int type1, int type2, size_t npts, const double *x,
                  double *f);

/* Genz test integrands over a box (family 0..5), for driver checks. */
int orc_genz(int family, unsigned dim, const double *a, const double *u,
             const double *xmin, const double *xmax, uint64_t maxeval, double reltol,
             orc_hcubature_v_fn hcub, double *val, double *err, orc_stats *stats);

/* Scalar primitives. */
void orc_sici(double x, double *si, double *ci);
double orc_aperture(double eta);
double orc_ncen(const double *hod, double m);
double orc_nsat(const double *hod, double m);
double orc_nsatnsat(const double *hod1, const double *hod2, int same, double A,
                    double epsilon, double m);
double orc_unfw(void *model, double k, double m, double z, double f);
double orc_bilin(void *model, double k1, double k2, double cosphi, double z);
double orc_fkernel(double k1, double k2, double cosphi);
/* Genz-Malik 7/5 rule and GK15 on one region; returns val, err, splitDim. */
int orc_rule_one_region(unsigned dim, orc_integrand_v f, void *fdata, const double *center,
                        const double *halfwidth, double *val, double *err, unsigned *split);
uint32_t orc_points_per_region(unsigned dim);

/* Sensitivity probe (tests only): a nonzero seed moves every exp, log, pow,
 * sin, cos and erf result inside the integrands by at most one unit in the
 * last place, by a hash of the argument, except the nearest-index chains
 * (masses 10^x and the logarithms that pick a table row); 0 restores the
 * exact functions. */
void orc_set_ulp_seed(uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
