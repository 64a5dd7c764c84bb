// souffle-gpu: command-line driver over the C ABI, mirroring the reference's run_simulation
// (src/md.cpp:29-693): the same flags and defaults (src/md.cpp:53-90), the template from
// infiles/<-f>, resume from the newest outfiles/restart_*.out (src/initialize.cpp:166-191), and
// the outfiles/ writers: energy.out (src/initialize.cpp:12-19, src/md.cpp:339-346, 616-623),
// model.parameters.out (src/md.cpp:112-180, 234), restart_<a>.out (src/md.cpp:526-536),
// ovito.lammpstrj (src/md.cpp:647-669) and the progress bar (src/functions.cpp:24-30).
// Encapsulation (-p, -a, -G, -m, -H): immobile big beads placed as src/initialize.cpp:566-680
// (or read from the newest outfiles/rb_*.out on a restart), capenergy.out rows
// (src/md.cpp:347-352, 625-630), rb_<a>.out beside each restart file (src/md.cpp:537-545) and
// the big beads appended to each movie frame (src/md.cpp:649-676).
// The step loop runs on the GPU in blocks between output steps.
//
// Initial state without a restart file: the reference's cubic lattice of identically oriented
// subunits (src/initialize.cpp:199-212) and its per-subunit uniform velocities
// (src/initialize.cpp:471-545). The reference draws those from GSL's default generator
// (mt19937) seeded with the first random() after srand(0) (src/rand_gen.h:18-25); std::mt19937
// with the same seed yields the same 32-bit stream and gsl_rng_uniform is x / 2^32.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "souffle_gpu.h"

namespace {

struct Opts {
  char response = 'm';
  std::string file = "41part_c";
  double C = 200, c = 500, S = 64, ks = 50, kb = 20, T = 2500, dt = 0.004, r = 1, damp = 100, q = 5, Q = 10,
         K = 298, e = 12, E = 2, L = 7.5, a = 3.0, G = -78, H = 3.5;
  bool R = true, X = false, V = true, p = false;
  int B = 20, M = 1000, W = 1000, w = 1000000, N = 4, m = 1;
  double skin = 0.3;       // GPU Verlet skin (no reference counterpart; -L is the reference's list)
  int fp32 = 0;            // --fp32 1: FP32 pair arithmetic
  double far_rf = 0;       // --far-rf: with -e 0, centre distance beyond which pairs are far field
                           // (0: 2.5 template diameters)
  double oligomers = 0;    // --oligomers <contact>: oligomer statistics at movie steps (0: off)
  int capsid = 20;         // --capsid <n>: subunits of a complete capsid (completion fraction)
  int digits = 17;         // --restart-digits <n>: significant digits of restart files (the
                           // reference writes 12, and 6 for the first velocity)
  int resume = 1;          // --resume <0|1>: continue from a restart's resume_<a>.txt sidecar when
                           // it exists (bit for bit); 0: restart as the reference does
};

bool parse_bool(const std::string& s) { return s == "1" || s == "true" || s == "yes" || s == "on"; }

int parse(int argc, char** argv, Opts& o) {
  std::map<std::string, std::string> kv;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      std::cout << "Usage:\nsouffle-gpu <options>  (flags of the reference: -D -f -C -c -S -s -b -T -t -r -d -q -Q -K -e\n"
                   "  -R -X -V -E -B -L -M -W -w -N -p -a -G -m -H; GPU only: --skin <sigma>, --fp32 <0|1>,\n"
                   "  --far-rf <sigma> (with -e 0), --oligomers <contact sigma>, --capsid <subunits>,\n"
                   "  --restart-digits <n>, --resume <0|1>)\n";
      return 1;
    }
    std::string key, val;
    if (a.rfind("--", 0) == 0) {
      key = a.substr(2);
      if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; return -1; }
      val = argv[++i];
    } else if (a.size() >= 2 && a[0] == '-') {
      key = a.substr(1, 1);
      if (a.size() > 2) val = a.substr(2);
      else {
        if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; return -1; }
        val = argv[++i];
      }
    } else {
      std::cerr << "unexpected argument " << a << "\n";
      return -1;
    }
    kv[key] = val;
  }
  for (auto& it : kv) {
    const std::string& k = it.first;
    const std::string& v = it.second;
    if (k == "D") o.response = v.empty() ? 'm' : v[0];
    else if (k == "f") o.file = v;
    else if (k == "C") o.C = atof(v.c_str());
    else if (k == "c") o.c = atof(v.c_str());
    else if (k == "S") o.S = atof(v.c_str());
    else if (k == "s") o.ks = atof(v.c_str());
    else if (k == "b") o.kb = atof(v.c_str());
    else if (k == "T") o.T = atof(v.c_str());
    else if (k == "t") o.dt = atof(v.c_str());
    else if (k == "r") o.r = atof(v.c_str());
    else if (k == "d") o.damp = atof(v.c_str());
    else if (k == "q") o.q = atof(v.c_str());
    else if (k == "Q") o.Q = atof(v.c_str());
    else if (k == "K") o.K = atof(v.c_str());
    else if (k == "e") o.e = atof(v.c_str());
    else if (k == "R") o.R = parse_bool(v);
    else if (k == "X") o.X = parse_bool(v);
    else if (k == "V") o.V = parse_bool(v);
    else if (k == "E") o.E = atof(v.c_str());
    else if (k == "B") o.B = atoi(v.c_str());
    else if (k == "L") o.L = atof(v.c_str());
    else if (k == "M") o.M = atoi(v.c_str());
    else if (k == "W") o.W = atoi(v.c_str());
    else if (k == "w") o.w = atoi(v.c_str());
    else if (k == "N") o.N = atoi(v.c_str());
    else if (k == "p") o.p = parse_bool(v);
    else if (k == "a") o.a = atof(v.c_str());
    else if (k == "G") o.G = atof(v.c_str());
    else if (k == "m") o.m = atoi(v.c_str());
    else if (k == "H") o.H = atof(v.c_str());
    else if (k == "skin") o.skin = atof(v.c_str());
    else if (k == "fp32") o.fp32 = atoi(v.c_str());
    else if (k == "far-rf") o.far_rf = atof(v.c_str());
    else if (k == "oligomers") o.oligomers = atof(v.c_str());
    else if (k == "capsid") o.capsid = atoi(v.c_str());
    else if (k == "restart-digits") o.digits = atoi(v.c_str());
    else if (k == "resume") o.resume = atoi(v.c_str());
    else { std::cerr << "unknown option -" << k << "\n"; return -1; }
  }
  return 0;
}

// The newest file in outfiles/ whose name contains `key` ("restart", or "rb" for the big
// beads), ordered by length then lexicographically (src/functions.cpp:262-307).
std::string newest_restart(const char* key = "restart") {
  std::vector<std::string> names;
  if (DIR* d = opendir("outfiles")) {
    while (dirent* e = readdir(d)) {
      std::string n = e->d_name;
      if (n.find(key) != std::string::npos) names.push_back(n);
    }
    closedir(d);
  }
  if (names.empty()) return "";
  std::sort(names.begin(), names.end(), [](const std::string& a, const std::string& b) {
    if (a.size() != b.size()) return a.size() < b.size();
    return a < b;
  });
  return names.back();
}

struct Template {
  std::vector<double> x, y, z, mass, q;
  std::vector<int> type;
  double si_sigma = 0, si_mass = 0, bondlength = 0;
};

bool read_template_beads(const std::string& path, Template& t) {
  std::ifstream in(path.c_str());
  if (!in) return false;
  std::string d;
  int np;
  in >> d >> d >> t.bondlength >> d >> d >> t.si_mass >> d >> d >> t.si_sigma;
  in >> d >> d >> d >> d >> d >> np;
  for (int i = 0; i < 8; i++) in >> d;
  for (int i = 0; i < np; i++) {
    long long name, type;
    double x, y, z, c, r, m;
    in >> name >> x >> y >> z >> type >> c >> r >> m;
    t.x.push_back(x); t.y.push_back(y); t.z.push_back(z);
    t.type.push_back((int)type); t.q.push_back(c); t.mass.push_back(m);
  }
  return (bool)in;
}

void progress(double fraction) {  // src/functions.cpp:24-30
  static const char* bar = "||||||||||||||||||||||||||||||||||||||||||||||||||||||||||||";
  int val = (int)(fraction * 100);
  int lpad = (int)(fraction * 60);
  int rpad = 60 - lpad;
  printf("\r%3d%% |%.*s%*s|", val, lpad, bar, rpad, "");
  fflush(stdout);
}

int fail(const char* what) {
  std::cerr << what << ": " << sf_last_error() << "\n";
  return 1;
}

// restart_<a>.out in the reference's layout (src/md.cpp:526-536): header, then per bead
// `i vx vy vz x y z`; every value at `digits` significant digits (17 round-trips a double)
void write_restart(const std::string& path, long long a, const std::vector<double>& pos, const std::vector<double>& vel,
                   long long N, int digits) {
  std::ofstream out(path.c_str(), std::ios::out);
  out << "Velocities & Positions for " << a << std::endl;
  out << std::setprecision(digits);
  for (long long i = 0; i < N; i++)
    out << i << "  " << vel[3 * i] << std::setw(25) << vel[3 * i + 1] << std::setw(25) << vel[3 * i + 2] << std::setw(25)
        << pos[3 * i] << std::setw(25) << pos[3 * i + 1] << std::setw(25) << pos[3 * i + 2] << std::setw(25) << std::endl;
}

// Sidecar of a restart: what the reference does not keep across a restart (the chain, the BD
// engine, the step counter and, for BD, the last integrator force) at full precision, so that a
// resumed run continues bit for bit.
void write_sidecar(const std::string& path, long long a, const sf_resume& r, const std::vector<double>* tforce) {
  std::ofstream out(path.c_str(), std::ios::out);
  out << std::setprecision(17);
  out << "iteration " << a << "\nstep " << r.step << "\nq_chain " << r.q_chain << "\nrng " << r.rng << "\nke " << r.ke
      << "\nxi";
  for (int j = 0; j < r.q_chain; j++) out << " " << r.xi[j];
  out << "\neta";
  for (int j = 0; j < r.q_chain; j++) out << " " << r.eta[j];
  out << "\ntforce " << (tforce ? tforce->size() / 3 : 0) << "\n";
  if (tforce)
    for (size_t i = 0; i < tforce->size(); i += 3) out << (*tforce)[i] << " " << (*tforce)[i + 1] << " " << (*tforce)[i + 2] << "\n";
}

bool read_sidecar(const std::string& path, long long& a, sf_resume& r, std::vector<double>& tforce) {
  std::ifstream in(path.c_str());
  if (!in) return false;
  std::string key;
  std::memset(&r, 0, sizeof(r));
  long long nt = 0;
  in >> key >> a >> key >> r.step >> key >> r.q_chain >> key >> r.rng >> key >> r.ke >> key;
  if (!in || r.q_chain < 0 || r.q_chain > 8) return false;
  for (int j = 0; j < r.q_chain; j++) in >> r.xi[j];
  in >> key;
  for (int j = 0; j < r.q_chain; j++) in >> r.eta[j];
  in >> key >> nt;
  tforce.assign(3 * (size_t)nt, 0.0);
  for (size_t i = 0; i < tforce.size(); i++) in >> tforce[i];
  return (bool)in;
}

}  // namespace

int main(int argc, char** argv) {
  Opts o;
  int pr = parse(argc, argv, o);
  if (pr != 0) return pr < 0 ? 1 : 0;
  if (o.X) { std::cerr << "cluster start (-X) needs infiles/xyz_T4, which this engine does not read\n"; return 1; }
  mkdir("outfiles", 0755);
  const bool brownian = o.response == 'b';
  std::ofstream traj("outfiles/energy.out", std::ios_base::app);
  std::ofstream encapdata("outfiles/capenergy.out", std::ios::out);
  std::ofstream ofile("outfiles/ovito.lammpstrj", std::ios_base::app);
  std::ofstream sysdata("outfiles/model.parameters.out", std::ios::out);
  traj << "time" << std::setw(15) << "kinetic" << std::setw(15) << "stretching" << std::setw(15) << "bending"
       << std::setw(15) << "LJ" << std::setw(15) << "electrostatics" << std::setw(15) << "total" << std::setw(15)
       << "potential" << std::setw(15) << "temperature" << std::setw(20) << "thermo_potential" << std::setw(15)
       << "thermo_kinetic" << std::setw(15) << "system_velocity_magnitude" << std::endl;
  if (brownian) sysdata << "Running brownian dynamics." << std::endl;
  else if (o.response == 'm') sysdata << "Running molecular dynamics with Nose Hoover thermostat." << std::endl;

  const std::string tpath = "infiles/" + o.file;
  sf_template* tmpl = sf_template_load(tpath.c_str());
  if (!tmpl) {
    std::cerr << "ERR: FILE " << o.file << " NOT OPENED. Check directory and/or filename.";
    return 1;
  }
  Template tb;
  read_template_beads(tpath, tb);
  int np = 0;
  sf_template_info(tmpl, &np, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  const int ns = (int)o.S;
  const long long N = (long long)ns * np;

  sf_cli cli{};
  cli.C_uM = o.C; cli.c_mM = o.c; cli.ks = o.ks; cli.kb = o.kb; cli.elj_att = o.E; cli.ecut_c = o.e;
  cli.dt = o.dt; cli.K = o.K; cli.Q = o.Q; cli.damp = o.damp; cli.n_subunits = ns;
  cli.q_chain = (int)o.q; cli.brownian = brownian ? 1 : 0;
  sf_params prm;
  if (sf_params_derive(tmpl, &cli, &prm)) return fail("parameters");
  prm.skin = o.skin;
  prm.pair_fp32 = o.fp32;
  // -e 0: Yukawa over all charged pairs of the box, near field from lists and far field by
  // subunit pairs beyond far_rf (sf_create takes 0 as 2.5 template diameters)
  prm.far_rf = o.far_rf;

  // model.parameters.out (src/md.cpp:135-180)
  const double Boltzmann = 1.3806e-23, Avagadro = 6.022e23;
  const double SIenergy = o.K * Boltzmann;
  const double SItime = std::sqrt(tb.si_sigma * tb.si_sigma * tb.si_mass / SIenergy);
  const double totaltime = o.T * o.dt;
  const double UnitEnergy = SIenergy * SItime * SItime / (tb.si_mass * tb.si_sigma * (1e-9));
  sysdata << "Simulation will run for " << totaltime * SItime / (1e-9) << " nanoseconds with a "
          << o.dt * SItime / (1e-12) << " picosecond timestep." << std::endl;
  sysdata << "Capsomere concentration: " << o.C << " micromolar" << std::endl;
  sysdata << "Salt concentration: " << o.c << " millimolar" << std::endl;
  sysdata << "Stretching constant: " << o.ks << " KbT" << "( " << (o.ks * tb.si_mass / (SItime * SItime)) << " N/m)" << std::endl;
  sysdata << "Bending constant: " << o.kb << " KbT" << "( " << (o.kb * tb.si_mass * tb.si_sigma * tb.si_sigma / (SItime * SItime))
          << " J)" << std::endl;
  sysdata << "Bondlength between beads is " << tb.bondlength << " LJ reduced units, which is "
          << tb.bondlength * tb.si_sigma / (1e-9) << " nanometers." << std::endl;
  sysdata << "Mass of a bead is " << tb.si_mass << " kg." << std::endl;
  sysdata << "Diameter of a bead is " << tb.si_sigma / (1e-9) << " nanometers." << std::endl;
  sysdata << "Total number of subunits is " << o.S << std::endl;
  sysdata << "Temperature is " << o.K << " K" << " Which is " << prm.T << " in reduced units" << std::endl;
  sysdata << "Attractive LJ paramerter is " << o.E << " Which is " << o.E * (SIenergy * Avagadro / (1000 * UnitEnergy))
          << " kJ/mol." << std::endl;
  sysdata << "Box length is " << prm.box * tb.si_sigma / (1e-9) << " nanometers." << std::endl;
  sysdata << "Screening length is " << 1 / prm.kappa << " nanometers." << std::endl;
  sysdata << "electrostatic cutoff is " << prm.ecut_el * tb.si_sigma / (1e-9) << " nanometers." << std::endl;
  double nlc = (prm.ecut_el > prm.ecut ? prm.ecut_el : prm.ecut) + o.L;
  sysdata << "Neighborlist cutoff is " << nlc * tb.si_sigma / (1e-9) << " nanometers." << std::endl;
  sysdata << "GPU Verlet skin is " << o.skin << " sigma (lists rebuilt on the device when a bead moves more than skin/2)."
          << std::endl;

  // initial state: restart file or the reference's lattice and velocities
  std::vector<double> pos(3 * N), vel(3 * N);
  long long loopStart = 0;
  sf_resume resume_rec;
  std::vector<double> resume_force;
  bool have_resume = false;
  bool restart = o.R;
  std::string rname = restart ? newest_restart() : "";
  if (rname.empty()) {
    if (restart) std::cout << "No restart files found! Starting simulation from scratch." << std::endl;
    restart = false;
  }
  if (restart) {
    std::cout << "Restarting from: " << rname << std::endl;
    std::ifstream rin(("outfiles/" + rname).c_str());
    if (!rin) { std::cerr << "ERR: RESTART FILE outfiles/restart.out NOT OPENED. Check directory and/or filename."; return 1; }
    std::string d;
    rin >> d >> d >> d >> d >> loopStart;
    for (long long i = 0; i < N; i++) rin >> d >> vel[3 * i] >> vel[3 * i + 1] >> vel[3 * i + 2] >> pos[3 * i] >> pos[3 * i + 1] >> pos[3 * i + 2];
    if (!rin) { std::cerr << "restart file has fewer beads than the system\n"; return 1; }
    // a sidecar written with this restart continues the run bit for bit from the next iteration
    if (o.resume) {
      std::ostringstream sn;
      sn << "outfiles/resume_" << loopStart << ".txt";
      long long a_side = -1;
      if (read_sidecar(sn.str(), a_side, resume_rec, resume_force)) {
        if (a_side != loopStart) { std::cerr << sn.str() << " belongs to another iteration\n"; return 1; }
        if (brownian && resume_force.size() != 3 * (size_t)N) { std::cerr << sn.str() << " lacks the BD force\n"; return 1; }
        have_resume = true;
        std::cout << "Resuming bit for bit from " << sn.str() << std::endl;
      }
    }
  } else {
    const long double bx = prm.box;
    const int nf = (int)std::ceil(std::pow((double)ns, 1.0 / 3.0));
    long long idx = 0;
    for (int i = 0; i < nf; i++)
      for (int j = 0; j < nf; j++)
        for (int k = 0; k < nf; k++) {
          if (idx >= ns) continue;
          for (int l = 0; l < np; l++) {
            long long g = idx * np + l;
            pos[3 * g] = (double)(((double)i * bx * (1 / (double)nf)) + (long double)tb.x[l]);
            pos[3 * g + 1] = (double)(((double)j * bx * (1 / (double)nf)) + (long double)tb.y[l]);
            pos[3 * g + 2] = (double)(((double)k * bx * (1 / (double)nf)) + (long double)tb.z[l]);
          }
          idx++;
        }
    srand(0);
    std::mt19937 gen((unsigned long)random());
    for (int s = 0; s < ns; s++) {
      double u[3];
      for (int a = 0; a < 3; a++) {
        double rn = gen() / 4294967296.0;
        u[a] = 0.5 * (rn) + (-0.5) * (1 - rn);
      }
      for (int l = 0; l < np; l++)
        for (int a = 0; a < 3; a++) vel[3 * ((long long)s * np + l) + a] = u[a];
    }
    long double av[3] = {0, 0, 0};
    for (long long i = 0; i < N; i++)
      for (int a = 0; a < 3; a++) av[a] += vel[3 * i + a];
    for (int a = 0; a < 3; a++) av[a] = av[a] * (1.0 / N);
    for (long long i = 0; i < N; i++)
      for (int a = 0; a < 3; a++) vel[3 * i + a] = (double)(vel[3 * i + a] - av[a]);
    double ke = 0;
    for (long long i = 0; i < N; i++) {
      long double v = std::sqrt((long double)vel[3 * i] * vel[3 * i] + (long double)vel[3 * i + 1] * vel[3 * i + 1] +
                                (long double)vel[3 * i + 2] * vel[3 * i + 2]);
      ke += (double)(0.5 * tb.mass[i % np] * v * v);
    }
    const double sc = std::sqrt(prm.T / (ke / (1.5 * N)));
    for (long long i = 0; i < 3 * N; i++) vel[i] = vel[i] * sc;
    std::ofstream lv("outfiles/list_initial_velocities.out", std::ios::out);
    for (int s = 0; s < ns; s++) {
      long long g = (long long)s * np;
      double vm = std::sqrt(vel[3 * g] * vel[3 * g] + vel[3 * g + 1] * vel[3 * g + 1] + vel[3 * g + 2] * vel[3 * g + 2]);
      lv << g << std::setw(15) << vel[3 * g] << "  " << vel[3 * g + 1] << "  " << vel[3 * g + 2] << "  " << vm << std::endl;
    }
  }

  sf_handle* h = sf_create(tmpl, 1, &ns, &prm);
  if (!h) return fail("sf_create");
  // encapsulation big beads (src/md.cpp:183-185, src/initialize.cpp:566-680): immobile, of
  // diameter 2a, charge G, mass 100 and type 3; read from the newest rb file on a restart,
  // placed among the subunit beads otherwise
  const int nbig = o.p ? o.m : 0;
  std::vector<double> big_pos(3 * (size_t)nbig), big_vel(3 * (size_t)nbig, 0.0);
  const double big_sigma = 2 * o.a, big_mass = 100.0;
  const int big_type = 3;
  double big_ke = 0;
  if (nbig > 0) {
    if (restart) {
      const std::string rb = newest_restart("rb");
      std::ifstream bin(("outfiles/" + rb).c_str());
      for (int i = 0; i < nbig; i++)
        bin >> big_vel[3 * i] >> big_vel[3 * i + 1] >> big_vel[3 * i + 2] >> big_pos[3 * i] >> big_pos[3 * i + 1] >>
            big_pos[3 * i + 2];
      if (rb.empty() || !bin) { std::cerr << "big-bead restart file (outfiles/rb_*.out) missing or short\n"; return 1; }
    } else if (sf_big_beads_place(tmpl, ns, pos.data(), prm.box, nbig, big_sigma, 1000000u, big_pos.data())) {
      std::cerr << "Error: " << sf_last_error() << std::endl;
      return 1;
    }
    for (int i = 0; i < nbig; i++) {  // BEAD::update_kinetic_energy (src/bead.h:129-131)
      const long double v = std::sqrt((long double)big_vel[3 * i] * big_vel[3 * i] +
                                      (long double)big_vel[3 * i + 1] * big_vel[3 * i + 1] +
                                      (long double)big_vel[3 * i + 2] * big_vel[3 * i + 2]);
      big_ke += (double)(0.5 * big_mass * v * v);
    }
    if (sf_set_big_beads(h, nbig, big_pos.data(), big_sigma, o.G, o.H)) return fail("sf_set_big_beads");
  }
  if (have_resume) {
    if (sf_set_state_resume(h, pos.data(), vel.data(), &resume_rec, resume_force.empty() ? nullptr : resume_force.data()))
      return fail("sf_set_state_resume");
    loopStart += 1;  // the restart's own iteration is done
  } else if (sf_set_state(h, pos.data(), vel.data())) {
    return fail("sf_set_state");
  }
  long long lib_step = have_resume ? resume_rec.step : 0;  // the library's loop-iteration counter
  // restart and sidecar of the current state (iteration a done)
  auto checkpoint = [&](const std::string& stem, long long a) -> int {
    if (sf_get_state(h, pos.data(), vel.data())) return fail("sf_get_state");
    sf_resume rec;
    std::vector<double> tf(brownian ? 3 * (size_t)N : 0);
    if (sf_get_resume(h, &rec, brownian ? tf.data() : nullptr)) return fail("sf_get_resume");
    std::ostringstream nm, sn;
    nm << "outfiles/" << stem << "_" << a << ".out";
    sn << "outfiles/" << (stem == "restart" ? "resume" : stem + "_resume") << "_" << a << ".txt";
    write_restart(nm.str(), a, pos, vel, N, o.digits);
    write_sidecar(sn.str(), a, rec, brownian ? &tf : nullptr);
    return 0;
  };
  std::vector<double> energies(7), per_bead(5 * N);
  double vcom = 0;
  auto energy_row = [&](long long a, bool with_vcom) -> int {
    if (sf_energies(h, energies.data(), per_bead.data())) return fail("sf_energies");
    const double ke = energies[0], se = energies[1], be = energies[2], ne = energies[3], ce = energies[4];
    const double tp = energies[5], tk = energies[6];
    const double tot = se + ke + ne + be + ce + tp + tk;
    traj << a << std::setw(15) << ke / N << std::setw(15) << se / N << std::setw(15) << be / N << std::setw(15) << ne / N
         << std::setw(15) << ce / N << std::setw(15) << tot / N << std::setw(15) << (be + se + ne + ce) / N << std::setw(15)
         << ke * 2 / (3 * N) << std::setw(15) << tp / N << std::setw(15) << tk / N;
    if (with_vcom) traj << std::setw(15) << vcom;
    traj << std::endl;
    if (nbig > 0) {  // src/md.cpp:347-352, 625-630
      double be2[2];
      if (sf_big_energies(h, be2)) return fail("sf_big_energies");
      const double bce = be2[0], bne = be2[1];
      encapdata << a << std::setw(15) << big_ke / nbig << std::setw(15) << bne / nbig << std::setw(15) << bce / nbig
                << std::setw(15) << (big_ke + bne + bce) / nbig << std::setw(15) << big_ke * 2 / (3 * nbig) << std::endl;
    }
    return 0;
  };
  if (energy_row(0, false)) return 1;
  // oligomer statistics (src/analysis.cpp:143-186, 267-285): one row per movie frame, and the
  // frame average of (oligomers of size k) x k written to analysis.ms at the end
  std::ofstream olig_out;
  std::vector<int> olig_label(ns), olig_ms(ns);
  std::vector<double> ms_weighted(ns, 0.0);
  long long olig_frames = 0;
  // completion fraction per movie frame in the layout of the reference's growth-kinetics table
  // (data/simulation/cluster.dat: label, time in ns, completion fraction, spread): here the
  // label is -b, the fraction (mean oligomer size in subunits - 1) / (capsid size - 1), which
  // is 0 for free subunits and 1 when every subunit sits in an oligomer of capsid size, and the
  // spread nan (one run has none)
  std::ofstream cluster_out;
  if (o.oligomers > 0) {
    if (o.capsid < 2) { std::cerr << "--capsid must be at least 2\n"; return 1; }
    olig_out.open("outfiles/oligomers.out", std::ios_base::app);
    olig_out << "time" << std::setw(15) << "oligomers" << std::setw(15) << "largest" << std::setw(15)
             << "mass_spectrum(1.." << ns << ")" << std::endl;
    cluster_out.open("outfiles/cluster.dat", std::ios_base::app);
    cluster_out << "# label  time(ns)  cluster_avg  std" << std::endl;
  }

  const double nsteps_d = totaltime / o.dt;
  long long a = loopStart;
  double percentagePre = -1;
  std::ofstream rb_out;
  while ((double)a < nsteps_d + 1) {
    // run up to the next iteration that writes output (or the last one)
    long long stop = a;
    while ((double)(stop + 1) < nsteps_d + 1 && stop % o.W != 0 && stop % o.w != 0 && stop % o.M != 0) stop++;
    const int rc_step = sf_step(h, (int)(stop - a + 1));
    if (rc_step == SF_E_NONFINITE) {
      // the handle holds the last finite state, after iteration a + k - 1 where k is how far the
      // library's step counter got in this call: kept for inspection under a name without
      // "restart", so that a rerun does not resume from it
      std::cerr << "\nsf_step: " << sf_last_error() << "\n";
      sf_resume rec;
      if (sf_get_resume(h, &rec, nullptr)) return fail("sf_get_resume");
      const long long a_last = a + (rec.step - lib_step) - 1;
      checkpoint("lastfinite", a_last);
      std::cerr << "last finite state (iteration " << a_last << ") written to outfiles/lastfinite_" << a_last << ".out\n";
      return 1;
    }
    if (rc_step) return fail("sf_step");
    lib_step += stop - a + 1;
    a = stop;
    if (a % o.w == 0) {  // src/md.cpp:526-536, with the sidecar of this build
      if (checkpoint("restart", a)) return 1;
      if (nbig > 0) {  // src/md.cpp:537-545
        std::ostringstream rn;
        rn << "outfiles/rb_" << a << ".out";
        rb_out.open(rn.str().c_str(), std::ios::out);
        rb_out << std::setprecision(o.digits);
        for (int i = 0; i < nbig; i++)
          rb_out << big_vel[3 * i] << std::setw(25) << big_vel[3 * i + 1] << std::setw(25) << big_vel[3 * i + 2]
                 << std::setw(25) << big_pos[3 * i] << std::setw(25) << big_pos[3 * i + 1] << std::setw(25)
                 << big_pos[3 * i + 2] << std::setw(25) << std::endl;
        rb_out.close();
      }
    }
    if (a % o.W == 0) {  // src/md.cpp:551-623
      sf_stats st;
      if (sf_get_stats(h, &st)) return fail("sf_get_stats");
      vcom = st.vcom;
      if (energy_row(a, true)) return 1;
    }
    if (o.oligomers > 0 && a % o.M == 0) {
      const int n = sf_oligomers(h, o.oligomers, olig_label.data(), olig_ms.data());
      if (n < 0) return fail("sf_oligomers");
      int largest = 0;
      for (int k = 0; k < ns; k++)
        if (olig_ms[k]) largest = k + 1;
      olig_out << a << std::setw(15) << n << std::setw(15) << largest;
      for (int k = 0; k < ns; k++) {
        olig_out << std::setw(8) << olig_ms[k];
        ms_weighted[k] += (double)olig_ms[k] * (k + 1);
      }
      olig_out << std::endl;
      olig_frames++;
      const double frac = ((double)ns / n - 1.0) / (o.capsid - 1.0);
      cluster_out << o.kb << "  " << std::fixed << std::setprecision(6) << a * o.dt * SItime / 1e-9 << "  " << frac
                  << "  nan" << std::defaultfloat << std::endl;
    }
    if (a % o.M == 0) {  // src/md.cpp:647-669
      sf_get_state(h, pos.data(), nullptr);
      ofile << "ITEM: TIMESTEP" << std::endl << a << std::endl << "ITEM: NUMBER OF ATOMS" << std::endl << N + nbig << std::endl
            << "ITEM: BOX BOUNDS" << std::endl << -prm.box / 2 << std::setw(15) << prm.box / 2 << std::endl << -prm.box / 2
            << std::setw(15) << prm.box / 2 << std::endl << -prm.box / 2 << std::setw(15) << prm.box / 2 << std::endl
            << "ITEM: ATOMS index type x y z b charge" << std::endl;
      for (long long b = 0; b < N; b++)
        ofile << b + 1 << std::setw(15) << tb.type[b % np] << std::setw(15) << pos[3 * b] << std::setw(15) << pos[3 * b + 1]
              << std::setw(15) << pos[3 * b + 2] << std::setw(15) << per_bead[5 * b + 1] << std::setw(15) << tb.q[b % np]
              << std::endl;
      for (int i = 0; i < nbig; i++)  // src/md.cpp:671-675 (a big bead's be stays 0)
        ofile << N + i + 1 << std::setw(15) << big_type << std::setw(15) << big_pos[3 * i] << std::setw(15)
              << big_pos[3 * i + 1] << std::setw(15) << big_pos[3 * i + 2] << std::setw(15) << 0.0 << std::setw(15) << o.G
              << std::endl;
    }
    const double percentage = std::round(a / (totaltime / o.dt) * 100 * 10) / 10;
    if (percentage != percentagePre) {
      progress(percentage / 100);
      percentagePre = percentage;
    }
    a++;
  }
  std::cout << std::endl;
  if (olig_frames > 0) {
    std::ofstream msdata("outfiles/analysis.ms");
    for (int k = 0; k < ns; k++) msdata << ms_weighted[k] / olig_frames << std::setw(15);
    msdata << std::endl;
  }
  sf_destroy(h);
  sf_template_free(tmpl);
  return 0;
}
