// ref_driver.cc -- a thin extern "C" driver over the reference's own hot-path
// headers, compiled from where they lie under the reference tree by
// oracle/Makefile into oracle/_ref/libsu2ref.so.
//
// TEST INFRASTRUCTURE, NOT PRODUCT.  Used by tests/ (to pin the C oracle to the
// reference's bits), by tests/golden/make_golden.py (fixtures) and by bench.py's
// cpu_baseline leg (the reference's own CPU path, timed).  Nothing here is
// reference source: it only includes the reference headers and calls them.
//
// dims[] = {Lx, Ly, Lz, Lt, nd}; g = 0 for U(1), 1 for SU(2).  Links and momenta
// are passed packed: SU(2) (Re a, Im a, Re b, Im b), U(1) phase; momenta
// (a,b,c) or (a).
#include "adjointfield.hh"
#include "exp_gauge.hh"
#include "flat-energy_density.hh"
#include "flat-gauge_energy.hpp"
#include "flat-gaugemonomial.hh"
#include "flat-gradient_flow.hh"
#include "gaugeconfig.hh"
#include "hamiltonian_field.hh"
#include "integrator.hh"
#include "md_params.hh"
#include "md_update.hh"
#include "monomial.hh"
#include "random_gauge_trafo.hh"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <random>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

template <class G> gaugeconfig<G> mk(const int *d, double beta = 0.0) {
  return gaugeconfig<G>(d[0], d[1], d[2], d[3], d[4], beta);
}

void put(gaugeconfig<_su2> &U, const double *src) {
  for (size_t i = 0; i < U.getSize(); ++i)
    U[i] = _su2(Complex(src[4 * i], src[4 * i + 1]), Complex(src[4 * i + 2], src[4 * i + 3]));
}
void put(gaugeconfig<_u1> &U, const double *src) {
  for (size_t i = 0; i < U.getSize(); ++i) U[i] = _u1(src[i]);
}
void get(const gaugeconfig<_su2> &U, double *dst) {
  for (size_t i = 0; i < U.getSize(); ++i) {
    _su2 u = U[i];
    dst[4 * i] = std::real(u.geta());
    dst[4 * i + 1] = std::imag(u.geta());
    dst[4 * i + 2] = std::real(u.getb());
    dst[4 * i + 3] = std::imag(u.getb());
  }
}
void get(const gaugeconfig<_u1> &U, double *dst) {
  for (size_t i = 0; i < U.getSize(); ++i) dst[i] = U[i].geta();
}
void putP(adjointfield<double, _su2> &P, const double *src) {
  for (size_t i = 0; i < P.getSize(); ++i) {
    P[i].seta(src[3 * i]);
    P[i].setb(src[3 * i + 1]);
    P[i].setc(src[3 * i + 2]);
  }
}
void putP(adjointfield<double, _u1> &P, const double *src) {
  for (size_t i = 0; i < P.getSize(); ++i) P[i].seta(src[i]);
}
void getP(const adjointfield<double, _su2> &P, double *dst) {
  for (size_t i = 0; i < P.getSize(); ++i) {
    dst[3 * i] = P[i].geta();
    dst[3 * i + 1] = P[i].getb();
    dst[3 * i + 2] = P[i].getc();
  }
}
void getP(const adjointfield<double, _u1> &P, double *dst) {
  for (size_t i = 0; i < P.getSize(); ++i) dst[i] = P[i].geta();
}

template <class G> void hotstart_t(const int *d, double *U, uint64_t seed, double delta) {
  auto C = mk<G>(d);
  hotstart(C, (int)seed, delta);
  get(C, U);
}

template <class G> void trafo_t(const int *d, double *U, uint64_t seed) {
  auto C = mk<G>(d);
  put(C, U);
  random_gauge_trafo(C, (int)seed);
  get(C, U);
}

template <class G>
void derivative_t(const int *d, const double *U, double beta, double xi, double fac, double *deriv) {
  auto C = mk<G>(d, beta);
  put(C, U);
  adjointfield<double, G> P(d[0], d[1], d[2], d[3], d[4]);
  adjointfield<double, G> D(d[0], d[1], d[2], d[3], d[4]);
  putP(D, deriv);
  hamiltonian_field<double, G> h(P, C);
  flat_spacetime::gaugemonomial<double, G> gm(0, xi);
  gm.derivative(D, h, fac);
  getP(D, deriv);
}

template <class G>
void integrate_t(const int *d, double *U, double *Pm, double beta, double xi, int kind, long nsteps,
                 double tau, int restore, long exponent) {
  auto C = mk<G>(d, beta);
  put(C, U);
  adjointfield<double, G> P(d[0], d[1], d[2], d[3], d[4]);
  putP(P, Pm);
  hamiltonian_field<double, G> h(P, C);
  kineticmonomial<double, G> km(0);
  km.setmdpassive();
  flat_spacetime::gaugemonomial<double, G> gm(0, xi);
  std::list<monomial<double, G> *> ml = {&km, &gm};
  integrator<double, G> *integ = set_integrator<double, G>((size_t)kind, (size_t)exponent);
  md_params mdp((size_t)nsteps, tau);
  integ->integrate(ml, h, mdp, restore != 0);
  delete integ;
  get(C, U);
  getP(P, Pm);
}

template <class G>
void md_update_t(const int *d, double *U, double beta, double xi, int kind, long nsteps, double tau,
                 long exponent, uint64_t seed, int revtest, double *out) {
  auto C = mk<G>(d, beta);
  put(C, U);
  kineticmonomial<double, G> km(0);
  km.setmdpassive();
  flat_spacetime::gaugemonomial<double, G> gm(0, xi);
  std::list<monomial<double, G> *> ml = {&km, &gm};
  integrator<double, G> *integ = set_integrator<double, G>((size_t)kind, (size_t)exponent);
  md_params mdp((size_t)nsteps, tau);
  if (revtest) mdp.enablerevtest();
  std::mt19937 engine(seed);
  md_update(C, engine, mdp, ml, *integ);
  delete integ;
  get(C, U);
  out[0] = mdp.getdeltaH();
  out[1] = mdp.getaccept();
  out[2] = revtest ? mdp.getdeltadeltaH() : NAN;
  out[3] = km.Hold;
  out[4] = gm.Hold;
  out[5] = km.Hnew;
  out[6] = gm.Hnew;
}

template <class G>
void hmc_run_t(const int *d, double *U, double beta, double xi, uint64_t seed, int heat, int kind,
               long nsteps, double tau, long exponent, long n_meas, long N_rev, double *rows) {
  // the flow of base_program::init_gauge_conf_mcmc and hmc_algo::run/do_hmc_step
  auto C = mk<G>(d, beta);
  const bool g_heat = heat != 0;
  hotstart(C, (int)seed, g_heat);
  random_gauge_trafo(C, 654321);
  const double fac = 2. / C.getndims() / (C.getndims() - 1);
  const double normalisation = fac / C.getVolume() / double(C.getNc());
  md_params mdparams((size_t)nsteps, tau);
  kineticmonomial<double, G> km(0);
  km.setmdpassive();
  flat_spacetime::gaugemonomial<double, G> gm(0, xi);
  std::list<monomial<double, G> *> ml = {&km, &gm};
  integrator<double, G> *integ = set_integrator<double, G>((size_t)kind, (size_t)exponent);
  double rate = 0.;
  for (long i = 0; i < n_meas; ++i) {
    mdparams.disablerevtest();
    if (i > 0 && N_rev != 0 && i % N_rev == 0) mdparams.enablerevtest();
    std::mt19937 engine(seed + i);
    md_update(C, engine, mdparams, ml, *integ);
    const double energy = flat_spacetime::gauge_energy(C);
    double E = 0., Q = 0.;
    flat_spacetime::energy_density(C, E, Q);
    rate += mdparams.getaccept();
    double *r = rows + 7 * i;
    r[0] = (double)i;
    r[1] = mdparams.getaccept();
    r[2] = energy * normalisation;
    r[3] = mdparams.getdeltaH();
    r[4] = rate / static_cast<double>(i + 1);
    r[5] = mdparams.getrevtest() ? mdparams.getdeltadeltaH() : NAN;
    r[6] = Q;
  }
  delete integ;
  get(C, U);
}

}  // namespace

REF_API void ref_mt_words(uint64_t seed, size_t n, uint32_t *out) {
  std::mt19937 e(seed);
  for (size_t i = 0; i < n; ++i) out[i] = (uint32_t)e();
}

REF_API void ref_normals(uint64_t seed, size_t n, double *out) {
  std::mt19937 e(seed);
  std::normal_distribution<double> normal(0., 1.);
  for (size_t i = 0; i < n; ++i) out[i] = normal(e);
}

REF_API void ref_uniforms(uint64_t seed, size_t n, double a, double b, double *out) {
  std::mt19937 e(seed);
  std::uniform_real_distribution<double> u(a, b);
  for (size_t i = 0; i < n; ++i) out[i] = u(e);
}

REF_API void ref_hotstart(int g, const int *d, double *U, uint64_t seed, double delta) {
  if (g == 1) hotstart_t<_su2>(d, U, seed, delta);
  else hotstart_t<_u1>(d, U, seed, delta);
}

REF_API void ref_random_gauge_trafo(int g, const int *d, double *U, uint64_t seed) {
  if (g == 1) trafo_t<_su2>(d, U, seed);
  else trafo_t<_u1>(d, U, seed);
}

REF_API void ref_initnormal(int g, const int *d, double *P, uint64_t seed) {
  std::mt19937 e(seed);
  if (g == 1) {
    adjointfield<double, _su2> A(d[0], d[1], d[2], d[3], d[4]);
    initnormal(e, A);
    getP(A, P);
  } else {
    adjointfield<double, _u1> A(d[0], d[1], d[2], d[3], d[4]);
    initnormal(e, A);
    getP(A, P);
  }
}

REF_API void ref_derivative(int g, const int *d, const double *U, double beta, double xi,
                            double fac, double *deriv) {
  if (g == 1) derivative_t<_su2>(d, U, beta, xi, fac, deriv);
  else derivative_t<_u1>(d, U, beta, xi, fac, deriv);
}

REF_API void ref_update_gauge(int g, const int *d, double *U, const double *Pm, double dtau) {
  if (g == 1) {
    auto C = mk<_su2>(d);
    put(C, U);
    adjointfield<double, _su2> P(d[0], d[1], d[2], d[3], d[4]);
    putP(P, Pm);
    hamiltonian_field<double, _su2> h(P, C);
    update_gauge(h, dtau);
    get(C, U);
  } else {
    auto C = mk<_u1>(d);
    put(C, U);
    adjointfield<double, _u1> P(d[0], d[1], d[2], d[3], d[4]);
    putP(P, Pm);
    hamiltonian_field<double, _u1> h(P, C);
    update_gauge(h, dtau);
    get(C, U);
  }
}

REF_API double ref_retr_sum_plaq(int g, const int *d, const double *U, double xi, int spatial_only) {
  const bool aniso = (xi != 1.0);
  if (g == 1) {
    auto C = mk<_su2>(d);
    put(C, U);
    return flat_spacetime::retr_sum_Wplaquettes(C, xi, aniso, spatial_only != 0);
  }
  auto C = mk<_u1>(d);
  put(C, U);
  return flat_spacetime::retr_sum_Wplaquettes(C, xi, aniso, spatial_only != 0);
}

REF_API double ref_kinetic(int g, const int *d, const double *Pm) {
  if (g == 1) {
    adjointfield<double, _su2> P(d[0], d[1], d[2], d[3], d[4]);
    putP(P, Pm);
    return 0.5 * (P * P);
  }
  adjointfield<double, _u1> P(d[0], d[1], d[2], d[3], d[4]);
  putP(P, Pm);
  return 0.5 * (P * P);
}

REF_API void ref_integrate(int g, const int *d, double *U, double *P, double beta, double xi,
                           int kind, long nsteps, double tau, int restore, long exponent) {
  if (g == 1) integrate_t<_su2>(d, U, P, beta, xi, kind, nsteps, tau, restore, exponent);
  else integrate_t<_u1>(d, U, P, beta, xi, kind, nsteps, tau, restore, exponent);
}

REF_API void ref_md_update(int g, const int *d, double *U, double beta, double xi, int kind,
                           long nsteps, double tau, long exponent, uint64_t seed, int revtest,
                           double *out) {
  if (g == 1) md_update_t<_su2>(d, U, beta, xi, kind, nsteps, tau, exponent, seed, revtest, out);
  else md_update_t<_u1>(d, U, beta, xi, kind, nsteps, tau, exponent, seed, revtest, out);
}

REF_API void ref_energy_density(int g, const int *d, const double *U, int clover, int ss,
                                double *E, double *Q) {
  if (g == 1) {
    auto C = mk<_su2>(d);
    put(C, U);
    flat_spacetime::energy_density(C, *E, *Q, clover != 0, ss != 0);
  } else {
    auto C = mk<_u1>(d);
    put(C, U);
    flat_spacetime::energy_density(C, *E, *Q, clover != 0, ss != 0);
  }
}

REF_API void ref_runge_kutta(int g, const int *d, double *U, double beta, double xi, double eps) {
  if (g == 1) {
    auto C = mk<_su2>(d, beta);
    put(C, U);
    adjointfield<double, _su2> Z(d[0], d[1], d[2], d[3], d[4]);
    hamiltonian_field<double, _su2> h(Z, C);
    flat_spacetime::gaugemonomial<double, _su2> SW(0, xi);
    runge_kutta(h, SW, eps);
    get(C, U);
  } else {
    auto C = mk<_u1>(d, beta);
    put(C, U);
    adjointfield<double, _u1> Z(d[0], d[1], d[2], d[3], d[4]);
    hamiltonian_field<double, _u1> h(Z, C);
    flat_spacetime::gaugemonomial<double, _u1> SW(0, xi);
    runge_kutta(h, SW, eps);
    get(C, U);
  }
}

// gradient_flow writes its table to `path` (flat-gradient_flow.hh:70-197).
REF_API void ref_gradient_flow(int g, const int *d, const double *U, double beta, double xi,
                               double tmax, double eps, const char *path) {
  if (g == 1) {
    auto C = mk<_su2>(d, beta);
    put(C, U);
    flat_spacetime::gradient_flow(C, std::string(path), tmax, eps, xi, 0.0, false);
  } else {
    auto C = mk<_u1>(d, beta);
    put(C, U);
    flat_spacetime::gradient_flow(C, std::string(path), tmax, eps, xi, 0.0, false);
  }
}

REF_API void ref_hmc_run(int g, const int *d, double *U, double beta, double xi, uint64_t seed,
                         int heat, int kind, long nsteps, double tau, long exponent, long n_meas,
                         long N_rev, double *rows) {
  if (g == 1) hmc_run_t<_su2>(d, U, beta, xi, seed, heat, kind, nsteps, tau, exponent, n_meas, N_rev, rows);
  else hmc_run_t<_u1>(d, U, beta, xi, seed, heat, kind, nsteps, tau, exponent, n_meas, N_rev, rows);
}

// raw-byte save (gaugeconfig.hh:225-230) of a packed SU(2)/U(1) field, for the
// file-format tests
REF_API void ref_save(int g, const int *d, const double *U, const char *path) {
  if (g == 1) {
    auto C = mk<_su2>(d);
    put(C, U);
    C.save(path);
  } else {
    auto C = mk<_u1>(d);
    put(C, U);
    C.save(path);
  }
}

REF_API int ref_load(int g, const int *d, double *U, const char *path) {
  if (g == 1) {
    auto C = mk<_su2>(d);
    int r = C.load(path);
    get(C, U);
    return r;
  }
  auto C = mk<_u1>(d);
  int r = C.load(path);
  get(C, U);
  return r;
}

REF_API void ref_su2_mul(const double *u, const double *v, double *out) {
  _su2 a(Complex(u[0], u[1]), Complex(u[2], u[3])), b(Complex(v[0], v[1]), Complex(v[2], v[3]));
  _su2 r = a * b;
  out[0] = std::real(r.geta());
  out[1] = std::imag(r.geta());
  out[2] = std::real(r.getb());
  out[3] = std::imag(r.getb());
}
