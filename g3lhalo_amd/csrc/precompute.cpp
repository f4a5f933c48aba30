// precompute.cpp -- the model's input tables from a cosmology and a source
// redshift distribution (g3l_precompute), the device counterpart of the
// reference's doPrecomputations.x (src/doPrecomputations.cpp:7-135).
//
// All adaptive integrals run as jobs of the batched 1-D Gauss-Kronrod engine
// (the n_bar kind, integrand modes of physics.h): first the sigma8
// normalisation and the comoving distances of the grid and of the source
// redshifts, then sigma^2(m, z) and d sigma^2/dm on the model grid and
// sigma^2 on the 256-mass grid of m*(z) -- about 4 N^2 integrals at N bins,
// each refined to the reference's 1e-4 relative tolerance.  The host keeps
// the scalar set-up (transfer-function parameters, growth, delta_c), the
// arg-min of m*(z) and the lensing-efficiency sums, in the reference's order;
// the closed-form 2-D tables are one device kernel.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "engine.h"

using namespace g3l;

namespace {

// std::numbers::pi, used by the reference's power-spectrum code
// (src/Powerspectrum.cpp:6,75,210-218,259); its cosmology and mass function
// use 3.14159 (src/constants.h:15).
constexpr double kPiExact = 3.141592653589793;
constexpr double kPiRef = 3.14159;

struct Cosmo {
  double Om, OL, Ob, s8, H0, ns;
  double c = 299792;          // km/s (src/Cosmology.h:26)
  double G = 1.327e11;        // km^3/Msun/s^2 (src/Cosmology.h:27)
  double Mpc_in_km = 3.086e19;  // (src/Cosmology.h:28)
  double rho_crit() const { return 3 * H0 * H0 * Mpc_in_km / 8 / kPiRef / G; }
  double E(double z) const { return Om * (1 + z) * (1 + z) * (1 + z) + OL; }
  double Om_z(double z) const { return Om * (1 + z) * (1 + z) * (1 + z) / E(z); }
  double delta_c(double z) const { return 1.686 * (1 + 0.0123 * std::log10(Om_z(z))); }
  double growth(double z) const {
    double om_m = Om_z(z);
    return 2.5 / (z + 1) * om_m /
           (std::pow(om_m, 4. / 7.) - OL + (1. + 0.5 * om_m) * (1 + OL / 70));
  }
};

// Eisenstein & Hu transfer-function parameters (src/Powerspectrum.cpp:15-69;
// the parameters the transfer function does not use are not kept).
void eh_params(const Cosmo &c, PrecompParams &p) {
  const double h = c.H0 / 100;
  p.omhh = c.Om * h * h;
  p.obhh = c.Ob * h * h;
  const double omhh = p.omhh, obhh = p.obhh;
  const double f_baryon = obhh / omhh;
  const double Tcmb = 2.72548;
  const double theta_cmb = Tcmb / 2.7;
  const double theta_cmbSq = theta_cmb * theta_cmb;
  const double theta_cmb4 = theta_cmbSq * theta_cmbSq;
  const double z_equality = 2.50e4 * omhh / theta_cmb4;
  p.k_equality = 0.0746 * omhh / theta_cmbSq;
  const double z_drag_b1 = 0.313 * std::pow(omhh, -0.419) * (1.0 + 0.607 * std::pow(omhh, 0.674));
  const double z_drag_b2 = 0.238 * std::pow(omhh, 0.223);
  const double z_drag = 1291.0 * std::pow(omhh, 0.251) / (1.0 + 0.659 * std::pow(omhh, 0.828)) *
                        (1.0 + z_drag_b1 * std::pow(obhh, z_drag_b2));
  const double R_drag = 31.5 * obhh / (theta_cmb4) * (1000.0 / (1.0 + z_drag));
  const double R_equality = 31.5 * obhh / (theta_cmb4) * (1000.0 / z_equality);
  p.sound_horizon = 2.0 / 3.0 / p.k_equality * std::sqrt(6.0 / R_equality) *
                    std::log((std::sqrt(1.0 + R_drag) + std::sqrt(R_drag + R_equality)) /
                             (1.0 + std::sqrt(R_equality)));
  p.k_silk = 1.6 * std::pow(obhh, 0.52) * std::pow(omhh, 0.73) *
             (1.0 + std::pow(10.4 * omhh, -0.95));
  const double alpha_c_a1 = std::pow(46.9 * omhh, 0.670) * (1.0 + std::pow(32.1 * omhh, -0.532));
  const double alpha_c_a2 = std::pow(12.0 * omhh, 0.424) * (1.0 + std::pow(45.0 * omhh, -0.582));
  p.alpha_c = std::pow(alpha_c_a1, -f_baryon) * std::pow(alpha_c_a2, -std::pow(f_baryon, 3));
  const double beta_c_b1 = 0.944 / (1.0 + std::pow(458.0 * omhh, -0.708));
  const double beta_c_b2 = std::pow(0.395 * omhh, -0.0266);
  p.beta_c = 1.0 / (1.0 + beta_c_b1 * (std::pow(1.0 - f_baryon, beta_c_b2) - 1.0));
  const double y = z_equality / (1.0 + z_drag);
  const double alpha_b_G =
      y * (-6.0 * std::sqrt(1.0 + y) +
           (2.0 + 3.0 * y) * std::log((std::sqrt(1.0 + y) + 1.0) / (std::sqrt(1.0 + y) - 1.0)));
  p.alpha_b = 2.07 * p.k_equality * p.sound_horizon * std::pow(1.0 + R_drag, -0.75) * alpha_b_G;
  p.beta_node = 8.41 * std::pow(omhh, 0.435);
  p.beta_b = 0.5 + f_baryon + (3.0 - 2.0 * f_baryon) * std::sqrt(std::pow(17.2 * omhh, 2.0) + 1.0);
}

// One 1-D job of the engine over [lo, hi] (the settings of every integral
// of the precomputation: tol 1e-4, no evaluation cap).
void add_job(EngineJobs &E, int mode, double lo, double hi, const double *aux, double t1,
             double t2, double t3) {
  JobDesc d{};
  d.kind = kNbar;
  d.dim = 1;
  d.pts = 15;
  d.args = (int)E.args.size();
  d.cap = 2048;
  d.reltol = 1e-4;
  d.maxeval = 0;
  TermArgs a{};
  a.mode = mode;
  a.aux = aux;
  a.t1 = t1;
  a.t2 = t2;
  a.t3 = t3;
  E.args.push_back(a);
  E.desc.push_back(d);
  double box[2 * kMaxDim] = {0};
  box[0] = lo;
  box[kMaxDim] = hi;
  E.boxes.insert(E.boxes.end(), box, box + 2 * kMaxDim);
}

// Runs the jobs in chunks that bound the region pool; values in job order.
int run_jobs(g3l_ctx *c, EngineJobs &all, std::vector<double> &val) {
  const size_t n = all.desc.size(), chunk = 16384;
  val.assign(n, 0.0);
  for (size_t s0 = 0; s0 < n; s0 += chunk) {
    const size_t s1 = std::min(n, s0 + chunk);
    EngineJobs E;
    E.desc.assign(all.desc.begin() + s0, all.desc.begin() + s1);
    E.args.assign(all.args.begin() + s0, all.args.begin() + s1);
    E.boxes.assign(all.boxes.begin() + s0 * 2 * kMaxDim, all.boxes.begin() + s1 * 2 * kMaxDim);
    for (auto &d : E.desc) d.args -= (int)s0;
    std::vector<JobState> st;
    if (int rc = run_engine(c, E, st, false, G3L_FP64)) return rc;
    for (size_t j = 0; j < st.size(); ++j) {
      if (st[j].flags & (kJobPoolFull | kJobNonfinite)) return status_of(st[j].flags);
      val[s0 + j] = st[j].val;
    }
  }
  return G3L_OK;
}

int precompute(g3l_ctx *c, const g3l_precompute_in &in, g3l_precompute_out &out) {
  const int N = in.nbins;
  const Cosmo cs{in.cosmo.Omega_m, in.cosmo.Omega_Lambda, in.cosmo.Omega_b, in.cosmo.sigma8,
                 in.cosmo.H0, in.cosmo.ns};
  const double rho_crit = cs.rho_crit();
  // grids (src/doPrecomputations.cpp:40-52)
  const double z_bin = (in.zmax - in.zmin) / N;
  const double m_bin = std::log(in.mmax / in.mmin) / N;
  const double k_bin = std::log(in.kmax / in.kmin) / N;
  std::vector<double> zs(N), ms(N), ks(N);
  for (int i = 0; i < N; ++i) {
    zs[i] = in.zmin + i * z_bin;
    ms[i] = std::exp(std::log(in.mmin) + m_bin * i);
    ks[i] = std::exp(std::log(in.kmin) + k_bin * i);
  }
  PrecompParams P{};
  P.Om = cs.Om;
  P.OL = cs.OL;
  P.H0 = cs.H0;
  P.c = cs.c;
  P.ns = cs.ns;
  eh_params(cs, P);
  DevBuf aux;
  if (int rc = aux.ensure(sizeof(PrecompParams))) return rc;
  const double *daux = aux.as<double>();
  G3L_HIP(hipMemcpy(aux.p, &P, sizeof(P), hipMemcpyHostToDevice));

  // phase 1: the sigma8 normalisation (src/Powerspectrum.cpp:71-80) and the
  // comoving distances of the grid and of the source redshifts
  // (src/Cosmology.cpp:91-102)
  const int Ns = in.nz_count;
  EngineJobs E1;
  add_job(E1, kOneDAmplitude, 0, 1e12, daux, 8. / cs.H0 * 100, 0, 0);
  for (int i = 0; i < N; ++i) add_job(E1, kOneDComoving, 0, zs[i], daux, 0, 0, 0);
  for (int i = 0; i < Ns; ++i) add_job(E1, kOneDComoving, 0, in.nz_z[i], daux, 0, 0, 0);
  std::vector<double> v1;
  if (int rc = run_jobs(c, E1, v1)) return rc;
  const double factor = kPiExact * cs.s8;
  P.amplitude = 2 * factor * factor / v1[0];
  G3L_HIP(hipMemcpy(aux.p, &P, sizeof(P), hipMemcpyHostToDevice));
  const double *wz = v1.data() + 1, *wsrc = v1.data() + 1 + N;

  // phase 2: sigma^2 and d sigma^2/dm on the model grid (src/HMF.cpp:7-31,
  // src/Powerspectrum.cpp:209-263) and sigma^2 on m*'s 256-mass grid
  // (src/Powerspectrum.cpp:110-126)
  const double D0 = cs.growth(0.0);
  constexpr int kNStar = 256;
  const double mstar_min = 1e10, mstar_max = 1e18;
  const double mstar_bin = std::log(mstar_max / mstar_min) / kNStar;
  std::vector<double> mgrid(kNStar);
  for (int s = 0; s < kNStar; ++s) mgrid[s] = std::exp(std::log(mstar_min) + mstar_bin * s);
  EngineJobs E2;
  std::vector<double> growth(N), delta_c(N), rho_mean(N);
  for (int i = 0; i < N; ++i) {
    const double z = zs[i];
    growth[i] = cs.growth(z) / D0;
    delta_c[i] = cs.delta_c(z);
    const double om = cs.Om_z(z);
    rho_mean[i] = om * rho_crit;
    const double rho_m = rho_crit * cs.Om_z(z);
    for (int j = 0; j < N; ++j) {
      const double m = ms[j];
      const double R = std::pow(0.75 / kPiExact * m / rho_crit / om, 1. / 3.);
      add_job(E2, kOneDSigma2, 0, 1e12, daux, R, growth[i], 0);
      const double Rd = 0.6204 * std::pow(m / rho_m, 1. / 3.);
      const double dR = 0.2068 * std::pow(rho_m * m * m, -1. / 3.);
      add_job(E2, kOneDDSigma2, 0, 1e12, daux, Rd, growth[i], dR);
    }
    for (int s = 0; s < kNStar; ++s) {
      const double R = std::pow(0.75 / kPiExact * mgrid[s] / rho_crit / om, 1. / 3.);
      add_job(E2, kOneDSigma2, 0, 1e12, daux, R, growth[i], 0);
    }
  }
  std::vector<double> v2;
  if (int rc = run_jobs(c, E2, v2)) return rc;
  const size_t per_z = 2 * (size_t)N + kNStar;
  std::vector<double> sig2((size_t)N * N), dsig2((size_t)N * N), mstar(N);
  for (int i = 0; i < N; ++i) {
    const double *v = v2.data() + i * per_z;
    for (int j = 0; j < N; ++j) {
      sig2[(size_t)i * N + j] = v[2 * j] / 2 / kPiExact / kPiExact;
      dsig2[(size_t)i * N + j] = v[2 * j + 1] / kPiExact / kPiExact;
    }
    // m*(z): the first grid mass with the smallest |delta_c - sigma|
    int best = 0;
    double bestd = 0;
    for (int s = 0; s < kNStar; ++s) {
      const double diff = delta_c[i] - std::sqrt(v[2 * N + s] / 2 / kPiExact / kPiExact);
      const double d = std::sqrt(diff * diff);
      if (s == 0 || d < bestd) {
        bestd = d;
        best = s;
      }
    }
    mstar[i] = mgrid[best];
  }

  // 1-D tables (src/Cosmology.cpp:104-151)
  for (int i = 0; i < N; ++i) {
    out.z[i] = zs[i];
    out.m[i] = ms[i];
    out.k[i] = ks[i];
    out.w[i] = wz[i];
    out.dwdz[i] = cs.c / cs.H0 / std::sqrt(cs.E(zs[i]));
    double g = 0;
    const double deltaZ = (in.nz_z[Ns - 1] - in.nz_z[0]) / Ns;
    for (int s = 0; s < Ns; ++s)
      if (in.nz_z[s] >= zs[i]) g += in.nz_n[s] * (wsrc[s] - wz[i]) / wsrc[s];
    out.g[i] = g * deltaZ;
  }

  // 2-D tables on the device
  const size_t n1 = N, n2 = (size_t)N * N;
  DevBuf dbuf;
  if (int rc = dbuf.ensure((7 * n1 + 6 * n2) * sizeof(double))) return rc;
  double *d = dbuf.as<double>();
  double *dzs = d, *dms = d + n1, *dks = d + 2 * n1, *dgr = d + 3 * n1, *ddc = d + 4 * n1,
         *drm = d + 5 * n1, *dmst = d + 6 * n1, *dsig = d + 7 * n1, *ddsig = dsig + n2,
         *dpl = ddsig + n2, *dco = dpl + n2, *dhm = dco + n2, *dbh = dhm + n2;
  hipStream_t st = c->stream;
  const std::vector<double> *ins[] = {&zs, &ms, &ks, &growth, &delta_c, &rho_mean, &mstar};
  for (int a = 0; a < 7; ++a)
    G3L_HIP(hipMemcpyAsync(d + a * n1, ins[a]->data(), n1 * sizeof(double), hipMemcpyHostToDevice,
                           st));
  G3L_HIP(hipMemcpyAsync(dsig, sig2.data(), n2 * sizeof(double), hipMemcpyHostToDevice, st));
  G3L_HIP(hipMemcpyAsync(ddsig, dsig2.data(), n2 * sizeof(double), hipMemcpyHostToDevice, st));
  PrecompTablesLaunch L{daux, dzs, dms, dks, dgr, ddc, drm, dmst, dsig, ddsig,
                        dpl, dco, dhm, dbh, N};
  G3L_HIP(launch_precomp_tables(L, st));
  G3L_HIP(hipMemcpyAsync(out.P_lin, dpl, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
  G3L_HIP(hipMemcpyAsync(out.concentration, dco, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
  G3L_HIP(hipMemcpyAsync(out.hmf, dhm, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
  G3L_HIP(hipMemcpyAsync(out.b_h, dbh, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
  G3L_HIP(hipStreamSynchronize(st));
  return G3L_OK;
}

}  // namespace

extern "C" int g3l_precompute(const g3l_precompute_in *in, g3l_precompute_out *out) {
  if (!in || !out || in->nbins < 1 || in->nz_count < 1 || !in->nz_z || !in->nz_n)
    return G3L_EINVAL;
  if (!out->z || !out->m || !out->k || !out->g || !out->w || !out->dwdz || !out->hmf ||
      !out->P_lin || !out->b_h || !out->concentration)
    return G3L_EINVAL;
  if (!(in->zmax > in->zmin) || !(in->mmin > 0) || !(in->mmax > in->mmin) || !(in->kmin > 0) ||
      !(in->kmax > in->kmin) || !(in->cosmo.H0 > 0) || !(in->cosmo.Omega_m > 0) ||
      !(in->cosmo.Omega_b > 0) || !(in->cosmo.Omega_b < in->cosmo.Omega_m))
    return G3L_EINVAL;
  g3l_ctx *c = nullptr;
  if (int rc = ctx_create_bare(&c)) return rc;
  int rc = precompute(c, *in, *out);
  g3l_ctx_destroy(c);
  return rc;
}
