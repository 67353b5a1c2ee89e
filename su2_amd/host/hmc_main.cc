// hmc_main.cc -- HMC host program over the GPU adapter (include/su2_mi355x/hmc_gpu.hh).
//
// Follows the reference's program flow for an HMC run:
//   run_program<G> (run_program.hpp:37-65) -> hmc_algo<G>::run (hmc.hpp:161-205):
//   init_gauge_conf_mcmc (base_program.hpp:276-318: hotstart or restart, then
//   random_gauge_trafo(654321)), md_params(n_steps, tau), monomials {kinetic
//   (md-passive), gauge(xi)}, set_integrator(name, exponent), and per trajectory
//   do_hmc_step (hmc.hpp:86-129: engine = mt19937(seed + i), md_update,
//   gauge_energy, energy_density, the output line) and after_MCMC_step
//   (base_program.hpp:439-456: save every N_save, nconf_counter.txt).
// Output lines and files keep the reference's formats (io.hh:41-62, :253-304).
//
// Input: the reference's YAML keys for this path (geometry, hmc, integrator,
// monomials.gauge) read by a small indentation parser (the reference links
// yaml-cpp, which is not part of this build), plus an optional
//   gpu:
//     precision: fp64 | fp32
//
//   usage: su2_hmc_gpu -f input.yaml [-g su2|u1]
//
// An input with a `metropolis:` section runs the Metropolis program instead
// (metropolis.hpp; keys n_meas, N_hit, delta, seed, heat, restart, N_save,
// conf_dir, conf_basename, lenghty_conf_name, beta_str_width).
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "su2_mi355x/hmc_gpu.hh"

namespace {

using Map = std::map<std::string, std::string>;

// "a.b.c" -> value for the indentation-nested "key: value" lines of a YAML file
Map parse_yaml_subset(const std::string &path) {
  std::ifstream in(path);
  if (!in) {
    std::cerr << "Error: could not open input file " << path << "\n";
    std::exit(1);
  }
  Map m;
  std::vector<std::pair<int, std::string>> stack;
  std::string line;
  while (std::getline(in, line)) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    const int indent = (int)line.find_first_not_of(' ');
    const auto colon = line.find(':');
    if (colon == std::string::npos) continue;
    auto trim = [](std::string s) {
      const auto a = s.find_first_not_of(" \t\r\"");
      const auto b = s.find_last_not_of(" \t\r\"");
      return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    const std::string key = trim(line.substr(0, colon));
    const std::string val = trim(line.substr(colon + 1));
    while (!stack.empty() && stack.back().first >= indent) stack.pop_back();
    std::string full;
    for (auto &p : stack) full += p.second + ".";
    full += key;
    if (val.empty()) stack.push_back({indent, key});
    else m[full] = val;
  }
  return m;
}

template <class T> T get(const Map &m, const std::string &k, T def) {
  auto it = m.find(k);
  if (it == m.end()) return def;
  std::istringstream ss(it->second);
  T v;
  ss >> v;
  return v;
}
bool getb(const Map &m, const std::string &k, bool def) {
  auto it = m.find(k);
  if (it == m.end()) return def;
  return it->second == "true" || it->second == "1";
}
std::string gets(const Map &m, const std::string &k, const std::string &def) {
  auto it = m.find(k);
  return it == m.end() ? def : it->second;
}

// io::read_nconf_counter (io.hh:267-287): second line "heat i path_conf"
void read_nconf_counter(const std::string &conf_dir, bool &heat, size_t &i, std::string &path) {
  std::ifstream nc(conf_dir + "nconf_counter.txt");
  if (!nc) {
    std::cerr << conf_dir + "nconf_counter.txt" << ": no such file or directory. Aborting.\n";
    std::abort();
  }
  std::string line;
  std::getline(nc, line);
  std::getline(nc, line);
  std::istringstream ls(line);
  double h = 0.0;
  ls >> h >> i >> path;
  heat = h != 0.0;
}

// io::update_nconf_counter (io.hh:289-304)
void write_nconf_counter(const std::string &conf_dir, bool heat, size_t i, const std::string &path) {
  std::ofstream nc(conf_dir + "nconf_counter.txt", std::ios::out);
  nc << "heat i path_conf\n" << double(heat) << " " << i << " " << path;
}

template <class G> int run(const Map &in) {
  namespace S = su2_mi355x;
  size_t Lx, Ly, Lz;
  if (in.count("geometry.L")) {
    Lx = Ly = Lz = get<size_t>(in, "geometry.L", 0);
  } else {
    Lx = get<size_t>(in, "geometry.X", 0);
    Ly = get<size_t>(in, "geometry.Y", 0);
    Lz = get<size_t>(in, "geometry.Z", 0);
  }
  const size_t Lt = get<size_t>(in, "geometry.T", 0);
  const size_t nd = get<size_t>(in, "geometry.ndims", 4);
  if (nd < 4) Lz = 1;  // parse_input_file.cc:57-67
  if (nd < 3) Ly = 1;
  const double beta = get<double>(in, "monomials.gauge.beta", 0.0);
  const double xi = get<double>(in, "monomials.gauge.anisotropic.xi", 1.0);
  const size_t n_meas = get<size_t>(in, "hmc.n_meas", 10);
  const size_t N_save = get<size_t>(in, "hmc.n_save", 100);
  const size_t seed = get<size_t>(in, "hmc.seed", 13526463);
  const bool heat = getb(in, "hmc.heat", true);
  const bool restart = getb(in, "hmc.restart", false);
  std::string conf_dir = gets(in, "hmc.conf_dir", ".");
  const std::string basename = gets(in, "hmc.conf_basename", "conf");
  const bool lengthy = getb(in, "hmc.lenghty_conf_name", true);
  const int bw = get<int>(in, "hmc.beta_str_width", 6);
  const std::string integ_name = gets(in, "integrator.name", "leapfrog");
  const size_t N_rev = get<size_t>(in, "integrator.N_rev", 0);
  const size_t n_steps = get<size_t>(in, "integrator.n_steps", 1000);
  const double tau = get<double>(in, "integrator.tau", 1.0);
  const size_t exponent = get<size_t>(in, "integrator.exponent", 0);
  const int prec = gets(in, "gpu.precision", "fp64") == "fp32" ? SU2GPU_FP32 : SU2GPU_FP64;
  // online gradient-flow measurement (omeasurements.hpp:65-91 via base_program.hpp:408-413)
  const bool do_omeas = getb(in, "hmc.do_omeas", false) || in.count("omeas.gradient_flow.tmax");
  const std::string res_dir = gets(in, "omeas.res_dir", ".");
  const size_t om_icounter = get<size_t>(in, "omeas.icounter", 0);
  const size_t om_nstep = get<size_t>(in, "omeas.nstep", 1);
  const size_t om_nmeas = get<size_t>(in, "omeas.n_meas", 1000000000);
  const double gf_eps = get<double>(in, "omeas.gradient_flow.epsilon", 0.01);
  const double gf_tmax = get<double>(in, "omeas.gradient_flow.tmax", 0.0);
  if (do_omeas && gf_tmax > 0) mkdir(res_dir.c_str(), 0755);

  std::cout << "## HMC Algorithm on MI355X (gfx950)\n";
  mkdir(conf_dir.c_str(), 0755);
  // io::get_conf_path_basename (io.hh:41-62)
  std::stringstream ss;
  ss << basename;
  if (lengthy) {
    ss << "." << Lx << "." << Ly << "." << Lz << "." << Lt << ".b" << std::fixed
       << std::setprecision(bw) << beta;
    if (xi != 1.0) ss << ".x" << std::fixed << std::setprecision(bw) << xi;
  }
  const std::string conf_path_basename = conf_dir + "/" + ss.str();

  S::gaugeconfig<G> U(Lx, Ly, Lz, Lt, nd, beta, xi, prec);
  size_t icounter = 0;
  bool g_heat = heat;
  if (restart) {
    std::string path;
    read_nconf_counter(conf_dir, g_heat, icounter, path);
    if (U.load(path) != 0) {
      std::cout << "Error: failed to load initial gauge configuration for intializing the "
                   "Markov chain Monte Carlo. Aborting.\n";
      std::abort();
    }
  } else {
    S::hotstart(U, (int)seed, g_heat);
  }
  const double fac = 2. / nd / (nd - 1);
  const double normalisation = fac / U.getVolume() / double(U.getNc());
  std::cout << "## Normalization factor: A = 2/(d*(d-1)*N_lat*N_c) = " << std::scientific
            << std::setw(18) << std::setprecision(15) << normalisation << "\n";
  std::cout << "## Acceptance rate parcentage: rho = rate/(i+1)\n";
  std::cout << "## Initial Plaquette: " << S::flat_spacetime::gauge_energy(U) * normalisation
            << std::endl;
  S::random_gauge_trafo(U, 654321);
  std::cout << "## Plaquette after rnd trafo: "
            << S::flat_spacetime::gauge_energy(U) * normalisation << std::endl;

  std::ofstream os(conf_dir + "/output.u1-hmc.data", icounter == 0 ? std::ios::out : std::ios::app);
  if (icounter == 0) {
    const std::string head = "i getaccept E*A dH rho ddH Q\n";
    std::cout << head;
    os << head;
  }
  S::md_params mdparams(n_steps, tau);
  S::kineticmonomial<double, G> km(0);
  km.setmdpassive();
  S::flat_spacetime::gaugemonomial<double, G> gm(0, xi);
  std::list<S::monomial<double, G> *> monomial_list = {&km, &gm};
  S::integrator<double, G> *md_integ = S::set_integrator<double, G>(integ_name, exponent);

  double rate = 0.;
  for (size_t i = icounter; i < n_meas + icounter; i++) {
    mdparams.disablerevtest();
    if (i > 0 && N_rev != 0 && i % N_rev == 0) mdparams.enablerevtest();
    std::mt19937 engine(seed + i);
    S::md_update(U, engine, mdparams, monomial_list, *md_integ);
    const double energy = S::flat_spacetime::gauge_energy(U);
    double E = 0., Q = 0.;
    S::flat_spacetime::energy_density(U, E, Q);
    rate += mdparams.getaccept();
    std::ostringstream line;
    line << i << " " << mdparams.getaccept() << " " << std::scientific << std::setw(18)
         << std::setprecision(15) << energy * normalisation << " " << std::setw(15)
         << mdparams.getdeltaH() << " " << std::setw(15) << rate / static_cast<double>(i + 1)
         << " ";
    if (mdparams.getrevtest()) line << mdparams.getdeltadeltaH();
    else line << "NA";
    line << " " << Q << "\n";
    std::cout << line.str() << std::flush;
    os << line.str() << std::flush;
    if (i > 0 && (i % N_save) == 0) {
      const std::string path_i = conf_path_basename + "." + std::to_string(i);
      U.save(path_i);
      // hmc.hpp:253-261 and base_program.hpp:367-413: flow the accepted
      // configuration, table in res_dir/gradient_flow.%06d
      const bool om = do_omeas && gf_tmax > 0 && i > om_icounter &&
                      (i - om_icounter) <= om_nmeas && (i % om_nstep) == 0 &&
                      mdparams.getaccept();
      if (om) {
        std::ostringstream gp;
        gp << res_dir << "/gradient_flow." << std::setw(6) << std::setfill('0') << i;
        S::flat_spacetime::gradient_flow(U, gp.str(), gf_tmax, gf_eps, xi, 0.0, false);
      }
      write_nconf_counter(conf_dir, g_heat, i, path_i);
    }
  }
  std::cout << "## Acceptance rate: " << rate / static_cast<double>(n_meas) << std::endl;
  U.save(conf_path_basename + ".final");
  delete md_integ;
  return 0;
}

// Metropolis program: metropolis<G>::run (metropolis.hpp:158-189) with
// do_sweep (:96-131) for one "thread": per sweep i, engine mt19937(seed + i),
// sweep(U, engines, delta, N_hit, beta, xi, anisotropic), then the line
// "i P E Q P_ss E_ss Q_ss" (energy_density without the clover definition),
// a configuration every N_save sweeps and the acceptance rates at the end.
template <class G> int run_metropolis(const Map &in) {
  namespace S = su2_mi355x;
  size_t Lx, Ly, Lz;
  if (in.count("geometry.L")) {
    Lx = Ly = Lz = get<size_t>(in, "geometry.L", 0);
  } else {
    Lx = get<size_t>(in, "geometry.X", 0);
    Ly = get<size_t>(in, "geometry.Y", 0);
    Lz = get<size_t>(in, "geometry.Z", 0);
  }
  const size_t Lt = get<size_t>(in, "geometry.T", 0);
  const size_t nd = get<size_t>(in, "geometry.ndims", 4);
  if (nd < 4) Lz = 1;
  if (nd < 3) Ly = 1;
  const double beta = get<double>(in, "monomials.gauge.beta", 0.0);
  const double xi = get<double>(in, "monomials.gauge.anisotropic.xi", 1.0);
  const bool aniso = xi != 1.0;
  const size_t n_meas = get<size_t>(in, "metropolis.n_meas", 10);
  const size_t N_save = get<size_t>(in, "metropolis.N_save", 100);
  const size_t seed = get<size_t>(in, "metropolis.seed", 13526463);
  const size_t N_hit = get<size_t>(in, "metropolis.N_hit", 10);
  const double delta = get<double>(in, "metropolis.delta", 1.0);
  const bool heat = getb(in, "metropolis.heat", true);
  const bool restart = getb(in, "metropolis.restart", false);
  std::string conf_dir = gets(in, "metropolis.conf_dir", ".");
  const std::string basename = gets(in, "metropolis.conf_basename", "conf");
  const bool lengthy = getb(in, "metropolis.lenghty_conf_name", true);
  const int bw = get<int>(in, "metropolis.beta_str_width", 6);
  const int prec = gets(in, "gpu.precision", "fp64") == "fp32" ? SU2GPU_FP32 : SU2GPU_FP64;

  std::cout << "## Metropolis Algorithm on MI355X (gfx950)\n";
  mkdir(conf_dir.c_str(), 0755);
  std::stringstream ss;
  ss << basename;
  if (lengthy) {
    ss << "." << Lx << "." << Ly << "." << Lz << "." << Lt << ".b" << std::fixed
       << std::setprecision(bw) << beta;
    if (xi != 1.0) ss << ".x" << std::fixed << std::setprecision(bw) << xi;
  }
  const std::string conf_path_basename = conf_dir + "/" + ss.str();

  S::gaugeconfig<G> U(Lx, Ly, Lz, Lt, nd, beta, xi, prec);
  // init_gauge_conf_mcmc (base_program.hpp:276-318): restart from
  // nconf_counter.txt or hot / cold start, then the gauge transformation
  size_t icounter = 0;
  bool g_heat = heat;
  if (restart) {
    std::cout << "## restart " << restart << "\n";
    std::string path;
    read_nconf_counter(conf_dir, g_heat, icounter, path);
    if (U.load(path) != 0) {
      std::cout << "Error: failed to load initial gauge configuration for intializing the "
                   "Markov chain Monte Carlo. Aborting.\n";
      std::abort();
    }
  } else {
    S::hotstart(U, (int)seed, g_heat);
  }
  const double normalisation = 2.0 / nd / (nd - 1) / U.getVolume() / double(U.getNc());
  std::cout << "## Initial Plaquette: " << S::flat_spacetime::gauge_energy(U) * normalisation
            << std::endl;
  S::random_gauge_trafo(U, 654321);
  std::ofstream os(conf_dir + "/output.u1-metropolis.data", icounter == 0 ? std::ios::out : std::ios::app);
  os << "## i P E Q P_ss E_ss Q_ss\n";
  const double facnorm = (nd > 2) ? double(nd) / double(nd - 2) : 0.0;
  double rate[2] = {0.0, 0.0};
  // run (metropolis.hpp:161-186) with one thread: sweep i = icounter .. icounter + n_meas
  for (size_t i = icounter; i < n_meas + icounter; ++i) {
    std::vector<std::mt19937> engines(1);
    engines[0].seed(seed + i);
    const std::vector<double> r = S::flat_spacetime::sweep(U, engines, delta, N_hit, beta, xi, aniso);
    rate[0] += r[0];
    rate[1] += r[1];
    std::ostringstream line;
    line << i;
    for (bool sp : {false, true}) {
      double E = 0., Q = 0.;
      S::flat_spacetime::energy_density(U, E, Q, false, sp);
      const double energy = S::flat_spacetime::gauge_energy(U, sp);
      line << " " << std::scientific << std::setprecision(15)
           << energy * normalisation * (sp ? facnorm : 1.0) << " " << E << " " << Q;
    }
    line << "\n";
    std::cout << line.str() << std::flush;
    os << line.str() << std::flush;
    // do_sweep / after_MCMC_step (metropolis.hpp:124-128, base_program.hpp:439-456)
    if (i > 0 && (i % N_save) == 0) {
      const std::string path_i = conf_path_basename + "." + std::to_string(i);
      U.save(path_i);
      write_nconf_counter(conf_dir, g_heat, i, path_i);
    }
  }
  std::cout << "## Acceptance rate " << rate[0] / double(n_meas) << " temporal acceptance rate "
            << rate[1] / double(n_meas) << std::endl;
  U.save(conf_path_basename + ".final");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::string input, group = "su2";
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if ((a == "-f" || a == "--file") && i + 1 < argc) input = argv[++i];
    else if ((a == "-g" || a == "--group") && i + 1 < argc) group = argv[++i];
  }
  if (input.empty()) {
    std::cout << "Error: invalid command line parameters.\nusage: su2_hmc_gpu -f input.yaml [-g su2|u1]\n";
    return 1;
  }
  std::cout << "## Parsing input file: " << input << "\n";
  const Map in = parse_yaml_subset(input);
  bool metropolis = false;
  for (const auto &kv : in)
    if (kv.first.rfind("metropolis.", 0) == 0) metropolis = true;
  try {
    if (metropolis) {
      if (group == "u1") return run_metropolis<su2_mi355x::_u1>(in);
      return run_metropolis<su2_mi355x::_su2>(in);
    }
    if (group == "u1") return run<su2_mi355x::_u1>(in);
    return run<su2_mi355x::_su2>(in);
  } catch (const std::exception &e) {
    std::cerr << "FATAL: " << e.what() << "\n";
    return 2;
  }
}
