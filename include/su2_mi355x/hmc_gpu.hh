// hmc_gpu.hh -- C++ adapter over the C ABI (include/su2_mi355x.h) with the
// names, signatures and call order of the reference's HMC layer, so a host
// program written against the reference (hmc.hpp, md_update.hh, integrator.hh,
// monomial.hh, flat-gaugemonomial.hh, gaugeconfig.hh, md_params.hh) runs the
// molecular dynamics on the GPU by including this header instead.
//
// Differences from the reference types, all forced by keeping U and P resident
// in HBM for a whole trajectory:
//   * gaugeconfig<G> owns a device lattice handle; element access goes through
//     get_data()/set_data() (host copies in the reference AoS order) instead of
//     operator[];
//   * momenta are generated on the host with the caller's URNG exactly as
//     initnormal does (adjointfield.hh:240-258), then uploaded;
//   * the monomial list may hold the kinetic monomial and one gauge monomial
//     (the pure-gauge action of hmc.hpp:58-84); anything else is rejected.
// Errors from the C ABI throw su2_mi355x::gpu_error.
#pragma once

#include <cmath>
#include <cstddef>
#include <fstream>
#include <iostream>
#include <list>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../su2_mi355x.h"
#include "normals.hh"

namespace su2_mi355x {

struct gpu_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what) {
  if (rc != SU2GPU_OK)
    throw gpu_error(std::string(what) + ": " + su2gpu_last_error());
}

// group tags (su2.hh:11-70, u1.hh:11-60)
struct _su2 {
  static constexpr int id = SU2GPU_SU2;
  static constexpr size_t N_c = 2;
  static constexpr size_t ncomp = 4;
  static constexpr size_t nmom = 3;
};
struct _u1 {
  static constexpr int id = SU2GPU_U1;
  static constexpr size_t N_c = 1;
  static constexpr size_t ncomp = 1;
  static constexpr size_t nmom = 1;
};
using su2 = _su2;

namespace spacetime_lattice {
const size_t nd_max = 4;
inline double num_pLloops_half(const int &d) { return d * (d - 1) / 2.0; }
}  // namespace spacetime_lattice

// md_params (md_params.hh:3-109)
class md_params {
public:
  md_params() {}
  md_params(size_t _nsteps, double _tau, size_t n = 1000000000)
    : nsteps(_nsteps), n_prec(n), tau(_tau) {}
  size_t getnsteps() const { return nsteps; }
  double gettau() const { return tau; }
  void setnsteps(size_t n) { nsteps = n; }
  void settau(double t) { tau = t; }
  void enablerevtest() { revtest = true; }
  void disablerevtest() { revtest = false; }
  bool getrevtest() const { return revtest; }
  void setdeltaH(double h) { deltaH = h; }
  double getdeltaH() const { return deltaH; }
  void setdeltadeltaH(double h) { deltadeltaH = h; }
  double getdeltadeltaH() const { return deltadeltaH; }
  void setH(double h) { H = h; }
  double getH() const { return H; }
  void setaccept(bool a) { accept = a; }
  bool getaccept() const { return accept; }
  void disableacceptreject() { acceptreject = false; }
  void enableacceptreject() { acceptreject = true; }
  bool getacceptreject() const { return acceptreject; }
  void setn_prec(size_t n) { n_prec = n; }
  size_t getn_prec() const { return n_prec; }
  void setexponent(size_t k) { exponent = k; }
  size_t getexponent() const { return exponent; }

private:
  size_t nsteps = 0, n_prec = 1000000000, exponent = 16;
  double tau = 1.0, H = 0., deltaH = 0., deltadeltaH = 0.;
  bool revtest = false, accept = true, acceptreject = true;
};

// gaugeconfig<G> (gaugeconfig.hh:40-223), resident on the GPU
template <class G> class gaugeconfig {
public:
  using value_type = G;
  gaugeconfig(const size_t Lx, const size_t Ly, const size_t Lz, const size_t Lt,
              const size_t ndims = spacetime_lattice::nd_max, const double beta = 0,
              const double xi = 1.0, const int precision = SU2GPU_FP64)
    : Lx(Lx), Ly(Ly), Lz(Lz), Lt(Lt), ndims(ndims) {
    check(su2gpu_create(&h, G::id, Lx, Ly, Lz, Lt, ndims, beta, xi, precision), "su2gpu_create");
  }
  gaugeconfig(const gaugeconfig &) = delete;
  gaugeconfig &operator=(const gaugeconfig &) = delete;
  ~gaugeconfig() { su2gpu_destroy(h); }

  size_t getLx() const { return Lx; }
  size_t getLy() const { return Ly; }
  size_t getLz() const { return Lz; }
  size_t getLt() const { return Lt; }
  size_t getndims() const { return ndims; }
  size_t getVolume() const { return Lx * Ly * Lz * Lt; }
  size_t getSize() const { return getVolume() * ndims; }
  int getNc() const { return (int)G::N_c; }
  double getBeta() const {
    double b;
    check(su2gpu_get_beta(h, &b), "su2gpu_get_beta");
    return b;
  }
  void setBeta(double b) { check(su2gpu_set_beta(h, b), "su2gpu_set_beta"); }
  void restoreSU() { check(su2gpu_restoreSU(h), "su2gpu_restoreSU"); }
  // host copies, packed reference AoS (4 doubles per SU(2) link, 1 per U(1) link)
  std::vector<double> get_data() const {
    std::vector<double> v(getSize() * G::ncomp);
    check(su2gpu_get_links(h, v.data(), SU2GPU_LAYOUT_PACKED), "su2gpu_get_links");
    return v;
  }
  void set_data(const std::vector<double> &v) {
    if (v.size() != getSize() * G::ncomp) throw gpu_error("set_data: wrong size");
    check(su2gpu_set_links(h, v.data(), SU2GPU_LAYOUT_PACKED), "su2gpu_set_links");
  }
  // gaugeconfig::save / load (gaugeconfig.hh:225-252): the reference's raw bytes
  void save(std::string const &path) const { check(su2gpu_save(h, path.c_str()), "save"); }
  int load(std::string const &path) {
    std::cout << "## Reading config from file " << path << std::endl;
    if (su2gpu_load(h, path.c_str()) != SU2GPU_OK) {
      std::cerr << "Error: could not read file from " << path << std::endl;
      return 1;
    }
    return 0;
  }
  su2gpu_lattice *handle() const { return h; }

private:
  su2gpu_lattice *h = nullptr;
  size_t Lx, Ly, Lz, Lt, ndims;
};

// hotstart / coldstart / random_gauge_trafo (gaugeconfig.hh:254-302,
// random_gauge_trafo.hh:10-33): the reference's host streams, uploaded
template <class G> void hotstart(gaugeconfig<G> &U, const int seed, const double delta) {
  check(su2gpu_hotstart(U.handle(), (uint64_t)seed, delta), "hotstart");
}
template <class G> void hotstart(gaugeconfig<G> &U, const int seed, const bool &hot) {
  check(su2gpu_hotstart(U.handle(), (uint64_t)seed, hot ? 1.0 : 0.0), "hotstart");
}
template <class G> void coldstart(gaugeconfig<G> &U) {
  check(su2gpu_coldstart(U.handle()), "coldstart");
}
template <class G> void random_gauge_trafo(gaugeconfig<G> &U, const int seed) {
  check(su2gpu_random_gauge_trafo(U.handle(), (uint64_t)seed), "random_gauge_trafo");
}

// adjointfield<Float, G> (adjointfield.hh:95-198): a host array in the
// reference order; the device momenta live in the lattice handle
template <typename Float, class G> class adjointfield {
public:
  adjointfield(size_t Lx, size_t Ly, size_t Lz, size_t Lt, size_t ndims = 4)
    : data(Lx * Ly * Lz * Lt * ndims * G::nmom), n(Lx * Ly * Lz * Lt * ndims) {}
  size_t getSize() const { return n; }
  std::vector<double> data;

private:
  size_t n;
};

template <class URNG, typename Float, class G>
void initnormal(URNG &engine, adjointfield<Float, G> &A) {
  if constexpr (std::is_same<URNG, std::mt19937>::value && std::is_same<Float, double>::value) {
    // the same draws, two passes (normals.hh)
    std::vector<double> pairs;
    detail::polar_normals(engine, A.data.data(), A.data.size(), pairs);
  } else {
    std::normal_distribution<double> normal(0., 1.);
    for (auto &x : A.data) x = Float(normal(engine));
  }
}

template <typename Float, class G> struct hamiltonian_field {
  adjointfield<Float, G> *momenta;
  gaugeconfig<G> *U;
  hamiltonian_field(adjointfield<Float, G> &m, gaugeconfig<G> &u) : momenta(&m), U(&u) {}
};

// monomials (monomial.hh:6-47, flat-gaugemonomial.hh:83-138)
template <typename Float, class G> class monomial {
public:
  double Hold = 0., Hnew = 0.;
  monomial(unsigned int ts = 0) : timescale(ts) {}
  virtual ~monomial() {}
  virtual void heatbath(hamiltonian_field<Float, G> const &h) = 0;
  virtual void accept(hamiltonian_field<Float, G> const &h) = 0;
  virtual bool is_kinetic() const { return false; }
  void reset() { Hold = Hnew = 0.; }
  void setTimescale(unsigned int t) { timescale = t; }
  unsigned int getTimescale() const { return timescale; }
  double getDeltaH() const { return Hnew - Hold; }
  void setmdactive() { mdactive = true; }
  void setmdpassive() { mdactive = false; }
  bool getmdactive() const { return mdactive; }

private:
  unsigned int timescale;
  bool mdactive = true;
};

template <typename Float, class G> class kineticmonomial : public monomial<Float, G> {
public:
  kineticmonomial(unsigned int ts) : monomial<Float, G>(ts) {}
  void heatbath(hamiltonian_field<Float, G> const &h) override {
    check(su2gpu_set_momenta(h.U->handle(), h.momenta->data.data()), "set_momenta");
    double K;
    check(su2gpu_kinetic_energy(h.U->handle(), &K), "kinetic");
    this->Hold = K;
  }
  void accept(hamiltonian_field<Float, G> const &h) override {
    double K;
    check(su2gpu_kinetic_energy(h.U->handle(), &K), "kinetic");
    this->Hnew = K;
  }
  bool is_kinetic() const override { return true; }
};

namespace flat_spacetime {
template <typename Float, class G> class gaugemonomial : public monomial<Float, G> {
public:
  gaugemonomial(unsigned int ts, const double &xi = 1.0) : monomial<Float, G>(ts), xi(xi) {}
  void heatbath(hamiltonian_field<Float, G> const &h) override {
    double S;
    check(su2gpu_gauge_action(h.U->handle(), &S), "gauge_action");
    this->Hold = S;
  }
  void accept(hamiltonian_field<Float, G> const &h) override {
    double S;
    check(su2gpu_gauge_action(h.U->handle(), &S), "gauge_action");
    this->Hnew = S;
  }
  double getxi() const { return xi; }

private:
  double xi;
};

// gauge_energy (flat-gauge_energy.hpp:109-137)
template <class G> double gauge_energy(const gaugeconfig<G> &U, bool spatial_only = false) {
  double r;
  check(su2gpu_gauge_energy(U.handle(), spatial_only ? 1 : 0, &r), "gauge_energy");
  return r;
}
// retr_sum_Wplaquettes (flat-gauge_energy.hpp:31-94)
template <class G>
double retr_sum_Wplaquettes(const gaugeconfig<G> &U, const double &xi = 1.0,
                            const bool &anisotropic = false, const bool &spatial_only = false) {
  double r;
  check(su2gpu_retr_sum_plaquettes(U.handle(), xi, anisotropic, spatial_only, &r), "retr_sum");
  return r;
}
// energy_density (flat-energy_density.hh:57-164)
template <class G>
void energy_density(const gaugeconfig<G> &U, double &res, double &Q, bool cloverdef = true,
                    const bool &ss = false) {
  check(su2gpu_energy_density(U.handle(), cloverdef, ss, &res, &Q), "energy_density");
}
// gradient_flow (flat-gradient_flow.hh:70-197): flows a copy of U on the GPU and
// writes the reference's table (header when tstart == 0, rows in scientific
// notation with 16 digits) to `path`, appending when tstart != 0
template <class G>
void gradient_flow(const gaugeconfig<G> &U, std::string const &path, const double &tmax,
                   const double &eps, const double & /*xi: the lattice's own*/,
                   const double &tstart, const bool & /*save_config*/) {
  std::vector<double> rows((size_t)(tmax / eps + 4) * 9);
  size_t n = 0;
  check(su2gpu_gradient_flow(U.handle(), tmax, eps, tstart, rows.data(), rows.size() / 9, &n),
        "gradient_flow");
  std::ofstream ofs(path, tstart == 0.0 ? std::ios::out : std::ios::app);
  if (tstart == 0.0) ofs << "t xi P P_ss Ep Ep_ss Ec Ec_ss Q" << std::endl;
  ofs << std::scientific;
  ofs.precision(16);
  for (size_t k = 0; k < n; ++k) {
    const double *r = &rows[9 * k];
    for (int c = 0; c < 9; ++c) ofs << r[c] << (c < 8 ? " " : "\n");
  }
}
// sweep(U, engines, delta, N_hit, beta, xi, anisotropic) (flat-sweep.hh:64-164):
// N_hit Metropolis hits on every link.  The caller's engines (a copy, as the
// reference takes them by value; metropolis.hpp:96-104 seeds them with
// seed + i + thread per sweep i) supply the key of the device's Philox stream:
// two draws of engines[0] make a 64-bit key, so the chain is reproducible from
// the caller's seeds as the reference's is, though with other draws.  beta
// becomes the lattice's beta; xi and the anisotropy are the lattice's own.
// Returns {rate, rate_time} as the reference does.
template <class URNG, class G>
std::vector<double> sweep(gaugeconfig<G> &U, std::vector<URNG> engines, const double &delta,
                          const size_t &N_hit, const double &beta, const double & /*xi*/ = 1.0,
                          const bool & /*anisotropic*/ = false) {
  if (engines.empty()) throw gpu_error("sweep: no engine");
  const uint64_t hi = (uint64_t)(uint32_t)engines[0](), lo = (uint64_t)(uint32_t)engines[0]();
  if (U.getBeta() != beta) U.setBeta(beta);
  double rate = 0.0, rate_time = 0.0;
  check(su2gpu_sweep(U.handle(), (hi << 32) | lo, 0, delta, N_hit, &rate, &rate_time), "sweep");
  return {rate, rate_time};
}
}  // namespace flat_spacetime

// integrators (integrator.hh:15-312)
enum integrators { LEAPFROG = 0, LP_LEAPFROG = 1, OMF4 = 2, LP_OMF4 = 3, EULER = 4, RUTH = 5, OMF2 = 6 };

template <typename Float, class G> class integrator {
public:
  explicit integrator(int id, size_t exponent = 0) : id(id), exponent(exponent) {}
  virtual ~integrator() {}
  virtual void integrate(std::list<monomial<Float, G> *> &, hamiltonian_field<Float, G> &h,
                         md_params const &params, const bool restore = true) {
    check(su2gpu_set_momenta(h.U->handle(), h.momenta->data.data()), "set_momenta");
    check(su2gpu_integrate(h.U->handle(), id, params.getnsteps(), params.gettau(), restore,
                           exponent),
          "integrate");
    check(su2gpu_get_momenta(h.U->handle(), h.momenta->data.data()), "get_momenta");
  }
  int getid() const { return id; }
  size_t getexponent() const { return exponent; }

private:
  int id;
  size_t exponent;
};
template <typename F, class G> struct leapfrog : integrator<F, G> { leapfrog() : integrator<F, G>(LEAPFROG) {} };
template <typename F, class G> struct lp_leapfrog : integrator<F, G> {
  lp_leapfrog(size_t n = 16) : integrator<F, G>(LP_LEAPFROG, n) {}
};
template <typename F, class G> struct omf2 : integrator<F, G> { omf2() : integrator<F, G>(OMF2) {} };
template <typename F, class G> struct omf4 : integrator<F, G> { omf4() : integrator<F, G>(OMF4) {} };
template <typename F, class G> struct lp_omf4 : integrator<F, G> {
  lp_omf4(size_t n = 16) : integrator<F, G>(LP_OMF4, n) {}
};
template <typename F, class G> struct euler : integrator<F, G> { euler() : integrator<F, G>(EULER) {} };
template <typename F, class G> struct ruth : integrator<F, G> { ruth() : integrator<F, G>(RUTH) {} };

template <typename F, class G> integrator<F, G> *set_integrator(const size_t integs, const size_t exponent) {
  switch (integs) {
    case LEAPFROG: std::cout << "leapfrog" << std::endl; return new leapfrog<F, G>();
    case LP_LEAPFROG: std::cout << "lp_leapfrog" << std::endl; return new lp_leapfrog<F, G>(exponent);
    case OMF4: std::cout << "omf4" << std::endl; return new omf4<F, G>();
    case LP_OMF4: std::cout << "lp_omf4" << std::endl; return new lp_omf4<F, G>(exponent);
    case EULER: std::cout << "euler" << std::endl; return new euler<F, G>();
    case RUTH: std::cout << "ruth" << std::endl; return new ruth<F, G>();
    case OMF2: std::cout << "omf2" << std::endl; return new omf2<F, G>();
    default:
      std::cout << "Integrator does not match, using default" << std::endl;
      return new leapfrog<F, G>();
  }
}
template <typename F, class G>
integrator<F, G> *set_integrator(const std::string &name, const size_t exponent) {
  integrator<F, G> *integ;
  if (name == "leapfrog") integ = new leapfrog<F, G>();
  else if (name == "lp_leapfrog") integ = new lp_leapfrog<F, G>(exponent);
  else if (name == "omf4") integ = new omf4<F, G>();
  else if (name == "lp_omf4") integ = new lp_omf4<F, G>(exponent);
  else if (name == "euler") integ = new euler<F, G>();
  else if (name == "ruth") integ = new ruth<F, G>();
  else if (name == "omf2") integ = new omf2<F, G>();
  else {
    std::cout << "Integrator does not match, using default" << std::endl;
    integ = new leapfrog<F, G>();
  }
  std::cout << "## MD integrator: " << name << std::endl;
  return integ;
}

// md_update (md_update.hh:18-81): momenta from the caller's engine, the whole
// trajectory on the GPU, Metropolis with the same engine, restore on reject
template <class URNG, typename Float, class G>
void md_update(gaugeconfig<G> &U, URNG &engine, md_params &params,
               std::list<monomial<Float, G> *> &monomial_list, integrator<Float, G> &md_integ) {
  monomial<Float, G> *kin = nullptr, *gauge = nullptr;
  for (auto *m : monomial_list) {
    if (m->is_kinetic()) kin = m;
    else if (dynamic_cast<flat_spacetime::gaugemonomial<Float, G> *>(m)) gauge = m;
    else throw gpu_error("md_update: only the kinetic and gauge monomials run on the GPU");
  }
  adjointfield<Float, G> momenta(U.getLx(), U.getLy(), U.getLz(), U.getLt(), U.getndims());
  initnormal(engine, momenta);
  std::uniform_real_distribution<Float> uniform(0., 1.);
  su2gpu_md_result r;
  check(su2gpu_md_trajectory(U.handle(), momenta.data.data(), md_integ.getid(),
                             params.getnsteps(), params.gettau(), md_integ.getexponent(),
                             params.getrevtest() ? 1 : 0, &r),
        "md_trajectory");
  if (kin) { kin->Hold = r.kin_old; kin->Hnew = r.kin_new; }
  if (gauge) { gauge->Hold = r.gauge_old; gauge->Hnew = r.gauge_new; }
  const double delta_H = r.deltaH;
  params.setdeltaH(delta_H);
  params.setaccept(true);
  if (delta_H > 0) {
    if (uniform(engine) > std::exp(-delta_H)) params.setaccept(false);
  }
  if (params.getrevtest()) params.setdeltadeltaH(r.deltadeltaH);
  check(su2gpu_md_finish(U.handle(), params.getaccept() ? 1 : 0), "md_finish");
}

}  // namespace su2_mi355x
