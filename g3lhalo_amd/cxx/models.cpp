// NNMap_Model and NMapMap_Model of the C++ API: the reference's classes
// (src/NNMap_Model.cu:15-415, src/NMapMap_Model.cu:14-261) on top of the C ABI.
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "../../include/g3lhalo/NMapMap_Model.h"
#include "../../include/g3lhalo/NNMap_Model.h"

namespace {

const uint64_t kMaxEval = 5000000;  // src/NNMap_Model.cu:318,343,368

void check(const char *what, int rc) {
  if (rc == G3L_OK || rc == G3L_MAXEVAL_HIT) return;
  std::fprintf(stderr, "g3lhalo: %s failed with status %d\n", what, rc);
  std::exit(rc);
}

g3l_hod to_hod(const g3lhalo::HOD *h) {
  const g3lhalo::Params *p = h->params;
  return g3l_hod{p->f, p->alpha, p->mmin, p->sigma, p->mprime, p->beta};
}

g3l_ctx *make_ctx(const g3lhalo::Cosmology *cosmo, double zmin, double zmax, double kmin,
                  double kmax, double mmin, double mmax, int nb, double *g, double *p1,
                  double *p2, double *w, double *dwdz, double *hmf, double *plin, double *bh,
                  double *conc) {
  g3l_tables t{};
  t.nbins = nb;
  t.zmin = zmin;
  t.zmax = zmax;
  t.kmin = kmin;
  t.kmax = kmax;
  t.mmin = mmin;
  t.mmax = mmax;
  t.g = g;
  t.p_lens1 = p1;
  t.p_lens2 = p2;
  t.w = w;
  t.dwdz = dwdz;
  t.hmf = hmf;
  t.P_lin = plin;
  t.b_h = bh;
  t.concentration = conc;
  t.H0 = cosmo->H0;
  t.Omega_m = cosmo->Omega_m;
  t.Omega_Lambda = cosmo->Omega_Lambda;
  g3l_ctx *ctx = nullptr;
  check("g3l_ctx_create", g3l_ctx_create(&t, &ctx));
  return ctx;
}

g3l_request request(int stat, int term, int set, double t1, double t2, double t3, int type1,
                    int type2, double tol) {
  g3l_request r{};
  r.stat = stat;
  r.term = term;
  r.hodset = set;
  r.type1 = type1;
  r.type2 = type2;
  r.theta[0] = t1;
  r.theta[1] = t2;
  r.theta[2] = t3;
  r.reltol = tol;
  r.maxeval = kMaxEval;
  return r;
}

// Tolerances of the reference's terms (src/NNMap_Model.cu:318,343,368;
// src/NMapMap_Model.cu:186,211,234).
double nnmap_tol(int term) { return term == 3 ? 1e-4 : 1e-3; }
const double kNMapMapTol = 1e-3;

}  // namespace

// ------------------------------------------------------------------ NNMap
g3lhalo::NNMap_Model::NNMap_Model(Cosmology *cosmology_, const double &zmin_,
                                  const double &zmax_, const double &kmin_, const double &kmax_,
                                  const double &mmin_, const double &mmax_, const int &Nbins_,
                                  double *g_, double *p_lens1_, double *p_lens2_, double *w_,
                                  double *dwdz_, double *hmf_, double *P_lin_, double *b_h_,
                                  double *concentration_, HOD *hod1_, HOD *hod2_, double A_,
                                  double epsilon_)
    : cosmology{cosmology_}, zmin{zmin_}, zmax{zmax_}, kmin{kmin_}, kmax{kmax_}, mmin{mmin_},
      mmax{mmax_}, Nbins{Nbins_}, g{g_}, p_lens1{p_lens1_}, p_lens2{p_lens2_}, w{w_},
      dwdz{dwdz_}, hmf{hmf_}, P_lin{P_lin_}, b_h{b_h_}, concentration{concentration_},
      hod1{hod1_}, hod2{hod2_}, A{A_}, epsilon{epsilon_} {
  zmin_integral = zmin;
  zmax_integral = zmax - (zmax - zmin) / Nbins;
  mmin_integral = mmin;
  mmax_integral = std::exp(std::log(mmin) + std::log(mmax / mmin) * (Nbins - 1) / Nbins);
  ctx_ = make_ctx(cosmology, zmin, zmax, kmin, kmax, mmin, mmax, Nbins, g, p_lens1, p_lens2, w,
                  dwdz, hmf, P_lin, b_h, concentration);
  rho_bar = new double[Nbins];
  n_bar1 = new double[Nbins];
  n_bar2 = new double[Nbins];
  updateDensities();
}

g3lhalo::NNMap_Model::~NNMap_Model() {
  if (ctx_) g3l_ctx_destroy(ctx_);
  delete[] rho_bar;
  delete[] n_bar1;
  delete[] n_bar2;
}

void g3lhalo::NNMap_Model::updateHODs(HOD *hod1_, HOD *hod2_) {
  hod1 = hod1_;
  hod2 = hod2_;
  updateDensities();
}

void g3lhalo::NNMap_Model::updateDensities() {
  g3l_hod a = to_hod(hod1), b = to_hod(hod2);
  check("g3l_hodset_clear", g3l_hodset_clear(ctx_));
  int32_t sid = -1;
  check("g3l_hodset_create",
        g3l_hodset_create(ctx_, &a, &b, A, epsilon, hod1 == hod2 ? 1 : 0, &sid));
  set_ = sid;
  check("g3l_hodset_densities", g3l_hodset_densities(ctx_, set_, rho_bar, n_bar1, n_bar2));
}

double g3lhalo::NNMap_Model::term(int t, double th1, double th2, double th3, int type1,
                                  int type2) {
  g3l_request r = request(G3L_NNMAP, t, set_, th1, th2, th3, type1, type2, nnmap_tol(t));
  g3l_result out{};
  check("g3l_eval_plain", g3l_eval_plain(ctx_, &r, &out));
  return out.value;
}

double g3lhalo::NNMap_Model::NNMap_1h(const double &theta1, const double &theta2,
                                      const double &theta3, const int &type1, const int &type2) {
  return term(1, theta1, theta2, theta3, type1, type2);
}
double g3lhalo::NNMap_Model::NNMap_2h(const double &theta1, const double &theta2,
                                      const double &theta3, const int &type1, const int &type2) {
  return term(2, theta1, theta2, theta3, type1, type2);
}
double g3lhalo::NNMap_Model::NNMap_3h(const double &theta1, const double &theta2,
                                      const double &theta3, const int &type1, const int &type2) {
  return term(3, theta1, theta2, theta3, type1, type2);
}
double g3lhalo::NNMap_Model::NNMap(const double &theta1, const double &theta2,
                                   const double &theta3, const int &type1, const int &type2) {
  return NNMap_1h(theta1, theta2, theta3, type1, type2) +
         NNMap_2h(theta1, theta2, theta3, type1, type2) +
         NNMap_3h(theta1, theta2, theta3, type1, type2);
}

void g3lhalo::NNMap_Model::terms(const std::vector<std::array<double, 3>> &thetas, int type1,
                                 int type2, std::vector<std::array<double, 3>> &out) {
  std::vector<g3l_request> reqs;
  reqs.reserve(thetas.size() * 3);
  for (const auto &th : thetas)
    for (int t = 1; t <= 3; ++t)
      reqs.push_back(
          request(G3L_NNMAP, t, set_, th[0], th[1], th[2], type1, type2, nnmap_tol(t)));
  std::vector<g3l_result> res(reqs.size());
  if (!reqs.empty())
    check("g3l_eval_batch", g3l_eval_batch(ctx_, reqs.data(), reqs.size(), precision, res.data()));
  out.resize(thetas.size());
  for (size_t i = 0; i < thetas.size(); ++i)
    for (int t = 0; t < 3; ++t) {
      check("integral", res[3 * i + t].status);
      out[i][t] = res[3 * i + t].value;
    }
}

void g3lhalo::NNMap_Model::calculateAll(const std::vector<double> &thetas1,
                                        const std::vector<double> &thetas2,
                                        const std::vector<double> &thetas3, HOD *hod1_,
                                        HOD *hod2_, double, double,
                                        std::vector<double> &results) {
  updateHODs(hod1_, hod2_);
  const size_t n = thetas1.size();
  const int types[3][2] = {{1, 2}, {1, 1}, {2, 2}};
  if (!batched) {
    for (auto &ty : types)
      for (size_t i = 0; i < n; ++i)
        results.push_back(NNMap(thetas1.at(i), thetas2.at(i), thetas3.at(i), ty[0], ty[1]));
    return;
  }
  std::vector<g3l_request> reqs;
  for (auto &ty : types)
    for (size_t i = 0; i < n; ++i)
      for (int t = 1; t <= 3; ++t)
        reqs.push_back(request(G3L_NNMAP, t, set_, thetas1.at(i), thetas2.at(i), thetas3.at(i),
                               ty[0], ty[1], nnmap_tol(t)));
  std::vector<g3l_result> res(reqs.size());
  if (!reqs.empty())
    check("g3l_eval_batch", g3l_eval_batch(ctx_, reqs.data(), reqs.size(), precision, res.data()));
  for (size_t j = 0; j < res.size(); j += 3)
    results.push_back(res[j].value + res[j + 1].value + res[j + 2].value);
}

// --------------------------------------------------------------- NMapMap
g3lhalo::NMapMap_Model::NMapMap_Model(Cosmology *cosmology_, const double &zmin_,
                                      const double &zmax_, const double &kmin_,
                                      const double &kmax_, const double &mmin_,
                                      const double &mmax_, const int &Nbins_, double *g_,
                                      double *p_lens_, double *w_, double *dwdz_, double *hmf_,
                                      double *P_lin_, double *b_h_, double *concentration_,
                                      HOD *hod_)
    : cosmology{cosmology_}, zmin{zmin_}, zmax{zmax_}, kmin{kmin_}, kmax{kmax_}, mmin{mmin_},
      mmax{mmax_}, Nbins{Nbins_}, g{g_}, p_lens{p_lens_}, w{w_}, dwdz{dwdz_}, hmf{hmf_},
      P_lin{P_lin_}, b_h{b_h_}, concentration{concentration_}, hod{hod_} {
  zmin_integral = zmin;
  zmax_integral = zmax - (zmax - zmin) / Nbins;
  mmin_integral = mmin;
  mmax_integral = std::exp(std::log(mmin) + std::log(mmax / mmin) * (Nbins - 1) / Nbins);
  ctx_ = make_ctx(cosmology, zmin, zmax, kmin, kmax, mmin, mmax, Nbins, g, p_lens, p_lens, w,
                  dwdz, hmf, P_lin, b_h, concentration);
  rho_bar = new double[Nbins];
  n_bar = new double[Nbins];
  updateDensities();
}

g3lhalo::NMapMap_Model::~NMapMap_Model() {
  if (ctx_) g3l_ctx_destroy(ctx_);
  delete[] rho_bar;
  delete[] n_bar;
}

void g3lhalo::NMapMap_Model::updateHOD(HOD *hod_) {
  hod = hod_;
  updateDensities();
}

void g3lhalo::NMapMap_Model::updateDensities() {
  g3l_hod a = to_hod(hod);
  check("g3l_hodset_clear", g3l_hodset_clear(ctx_));
  int32_t sid = -1;
  check("g3l_hodset_create", g3l_hodset_create(ctx_, &a, &a, 0.0, 0.0, 1, &sid));
  set_ = sid;
  check("g3l_hodset_densities", g3l_hodset_densities(ctx_, set_, rho_bar, n_bar, nullptr));
}

double g3lhalo::NMapMap_Model::term(int t, double th1, double th2, double th3) {
  g3l_request r = request(G3L_NMAPMAP, t, set_, th1, th2, th3, 1, 1, kNMapMapTol);
  g3l_result out{};
  check("g3l_eval_plain", g3l_eval_plain(ctx_, &r, &out));
  return out.value;
}

double g3lhalo::NMapMap_Model::NMapMap_1h(const double &theta1, const double &theta2,
                                          const double &theta3) {
  return term(1, theta1, theta2, theta3);
}
double g3lhalo::NMapMap_Model::NMapMap_2h(const double &theta1, const double &theta2,
                                          const double &theta3) {
  return term(2, theta1, theta2, theta3);
}
double g3lhalo::NMapMap_Model::NMapMap_3h(const double &theta1, const double &theta2,
                                          const double &theta3) {
  return term(3, theta1, theta2, theta3);
}
double g3lhalo::NMapMap_Model::NMapMap(const double &theta1, const double &theta2,
                                       const double &theta3) {
  return NMapMap_1h(theta1, theta2, theta3) + NMapMap_2h(theta1, theta2, theta3) +
         NMapMap_3h(theta1, theta2, theta3);
}

void g3lhalo::NMapMap_Model::terms(const std::vector<std::array<double, 3>> &thetas,
                                   std::vector<std::array<double, 3>> &out) {
  std::vector<g3l_request> reqs;
  for (const auto &th : thetas)
    for (int t = 1; t <= 3; ++t)
      reqs.push_back(request(G3L_NMAPMAP, t, set_, th[0], th[1], th[2], 1, 1, kNMapMapTol));
  std::vector<g3l_result> res(reqs.size());
  if (!reqs.empty())
    check("g3l_eval_batch", g3l_eval_batch(ctx_, reqs.data(), reqs.size(), precision, res.data()));
  out.resize(thetas.size());
  for (size_t i = 0; i < thetas.size(); ++i)
    for (int t = 0; t < 3; ++t) {
      check("integral", res[3 * i + t].status);
      out[i][t] = res[3 * i + t].value;
    }
}

void g3lhalo::NMapMap_Model::calculateAll(const std::vector<double> &thetas1,
                                          const std::vector<double> &thetas2,
                                          const std::vector<double> &thetas3, HOD *hod_,
                                          std::vector<double> &results) {
  updateHOD(hod_);
  const size_t n = thetas1.size();
  if (!batched) {
    for (size_t i = 0; i < n; ++i)
      results.push_back(NMapMap(thetas1.at(i), thetas2.at(i), thetas3.at(i)));
    return;
  }
  std::vector<std::array<double, 3>> th(n), out;
  for (size_t i = 0; i < n; ++i) th[i] = {thetas1.at(i), thetas2.at(i), thetas3.at(i)};
  terms(th, out);
  for (size_t i = 0; i < n; ++i) results.push_back(out[i][0] + out[i][1] + out[i][2]);
}
