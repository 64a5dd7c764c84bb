// CPU oracle for the SOUFFLE force-and-integrate path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke check and
// the cpu_baseline leg of bench.py load it; the shipped GPU library never does.
//
// It restates, object by object, the reference's double / x87 long double CPU path:
//   - template reader            src/initialize.cpp:46-128
//   - edge/face association       src/initialize.cpp:348-404 (done per template: edges and
//                                  faces of different subunits never match, so the result is
//                                  the same as the reference's global O(F*E) scan)
//   - centre bead                 src/initialize.cpp:266-288
//   - unit/cutoff derivation      src/md.cpp:135-173, list capacity src/md.cpp:229-239
//   - forceCalculation            src/forces.cpp:13-171 (OpenMP dynamic over subunits)
//   - edge length / face normal   src/edge.cpp:35-49, src/face.cpp:11-17, src/face.cpp:27-34
//   - stretching force/energy     src/bead.cpp:9-23
//   - bending force/energy        src/edge.cpp:51-188 (including the in-place normal flip)
//   - neighbour list              src/functions.cpp:69-117 (centre-bead prefilter, stale r2)
//   - NH chain                    src/functions.cpp:33-58, src/thermostat.h:35-52
//   - pair energies               src/energies.cpp:13-93
//   - step loop body (NH and BD)  src/md.cpp:400-514, bead updates src/bead.h:72-149
//   - oligomers (mass spectrum)   src/analysis.cpp:143-186
//   - encapsulation big beads     placement src/initialize.cpp:566-680, forces
//                                  src/forces.cpp:174-236, energies src/energies.cpp:96-136,
//                                  their BD draws src/md.cpp:482-487
// Templated on the real type R: R = long double mirrors the reference's VECTOR3D
// (src/vector3d.h:22); R = double is the FP64 restatement the GPU path is compared with.
// Quantities the reference keeps in `double` (edge length, face area, bead energies,
// masses, charges, the bending S) stay double in both modes.
//
// The BD noise uses std::default_random_engine and std::uniform_real_distribution
// themselves (libstdc++), seeded 69521 as src/md.cpp:382-385.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

template <typename R>
struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  R dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
  V3 cross(const V3& o) const {
    return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
  }
  V3 scale(R s) const { return V3(x * s, y * s, z * s); }
  R mag2() const { return x * x + y * y + z * z; }
  R mag() const { return std::sqrt(x * x + y * y + z * z); }
  void operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; }
  void operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; }
};

// ---------------------------------------------------------------- template reader
// src/initialize.cpp:46-128: whitespace-token reader; three header lines, then the
// particle, edge, triangle and LJ-attraction sections. The repulsion section is unread.
struct Tmpl {
  double bondlength = 0, SImass = 0, SIsigma = 0;
  int np = 0, ne = 0, nf = 0, na = 0;
  std::vector<long double> x, y, z, q, sigma, mass;
  std::vector<int> type;
  std::vector<int> e1, e2, etype;
  std::vector<long double> len0;
  std::vector<int> t1, t2, t3, ftype;
  std::vector<int> la1, la2;  // ordered (type1, type2) rows of lj_a
};

bool read_template(const char* path, Tmpl& t) {
  std::ifstream in(path);
  if (!in) { g_err = std::string("cannot open template ") + path; return false; }
  std::string d;
  in >> d >> d >> t.bondlength >> d >> d >> t.SImass >> d >> d >> t.SIsigma;
  in >> d >> d >> d >> d >> d >> t.np;
  for (int i = 0; i < 8; i++) in >> d;
  for (int i = 0; i < t.np; i++) {
    unsigned name, type;
    long double x, y, z, c, r, m;
    in >> name >> x >> y >> z >> type >> c >> r >> m;
    t.x.push_back(x); t.y.push_back(y); t.z.push_back(z);
    t.type.push_back((int)type); t.q.push_back(c); t.sigma.push_back(r); t.mass.push_back(m);
  }
  in >> d >> d >> d >> d >> d >> t.ne;
  for (int i = 0; i < 5; i++) in >> d;
  for (int i = 0; i < t.ne; i++) {
    unsigned name, type; int a, b; long double len;
    in >> name >> a >> b >> type >> len;
    t.e1.push_back(a); t.e2.push_back(b); t.etype.push_back((int)type); t.len0.push_back(len);
  }
  in >> d >> d >> d >> d >> d >> t.nf;
  for (int i = 0; i < 6; i++) in >> d;
  for (int i = 0; i < t.nf; i++) {
    unsigned name, type; int a, b, c, normal;
    in >> name >> a >> b >> c >> type >> normal;
    t.t1.push_back(a); t.t2.push_back(b); t.t3.push_back(c); t.ftype.push_back((int)type);
  }
  in >> d >> d >> d >> d >> d >> d >> t.na;
  for (int i = 0; i < 5; i++) in >> d;
  for (int i = 0; i < t.na; i++) {
    unsigned name; int a, b; double eps, sig;
    in >> name >> a >> b >> eps >> sig;
    t.la1.push_back(a); t.la2.push_back(b);
  }
  if (!in || t.np <= 0 || t.ne < 0 || t.nf < 0) { g_err = "malformed template"; return false; }
  return true;
}

struct Thermo {  // src/thermostat.h:8-54
  double kB, Q, T; unsigned long dof; double xi, eta, pe, ke;
  Thermo(double Q_, double T_, unsigned long dof_) : kB(1), Q(Q_), T(T_), dof(dof_), xi(0), eta(0), pe(0), ke(0) {}
  void update_eta(double dt) { if (Q == 0) return; eta = eta + 0.5 * dt * xi; }
};

// src/functions.cpp:33-58
void update_chain_xi(unsigned j, std::vector<Thermo>& bath, double dt, long double ke) {
  if (bath[j].Q == 0) return;
  if (j != 0) {
    bath[j].xi = bath[j].xi * std::exp(-0.5 * dt * bath[j + 1].xi) +
                 0.5 * dt * (1.0 / bath[j].Q) *
                     (bath[j - 1].Q * bath[j - 1].xi * bath[j - 1].xi - bath[j].dof * bath[j].kB * bath[j].T) *
                     std::exp(-0.25 * dt * bath[j + 1].xi);
  } else {
    bath[j].xi = bath[j].xi * std::exp(-0.5 * dt * bath[j + 1].xi) +
                 0.5 * dt * (1.0 / bath[j].Q) * (2 * ke - bath[j].dof * bath[j].kB * bath[j].T) *
                     std::exp(-0.25 * dt * bath[j + 1].xi);
  }
}

struct Cli {  // the reference's flags (src/md.cpp:53-90) that reach the hot path
  double C_uM, c_mM, ks, kb, elj_att, ecut_c, L_skin, dt, K, Q, damp;
  int n_subunits, B_build, q_chain, brownian;
};

struct SysBase {
  virtual ~SysBase() {}
};

template <typename R>
struct Sys : SysBase {
  typedef V3<R> Vec;
  Tmpl t;
  Cli cli;
  int Ns = 0, N = 0;
  // derived parameters (src/md.cpp:135-173)
  double SItime, T, box_x, ecut, lb, ni, kappa, ecut_el, NListCutoff, qs = 1;
  unsigned NLV;
  Vec box, hbox;
  // beads
  std::vector<Vec> pos, vel, tforce, sforce, bforce, fdrag, fran;
  std::vector<double> m, sigma, q, se, ke, be, ne, ce;
  std::vector<int> type, unit;
  std::vector<std::vector<int>> itsE;  // incident edges, ascending edge id
  std::vector<std::vector<int>> itsN;  // neighbour list, -1 terminated (fixed capacity)
  // edges
  std::vector<int> eb0, eb1, etype;
  std::vector<double> elen0, elength;
  std::vector<Vec> elenvec;
  std::vector<std::vector<int>> eF;  // faces of the edge, ascending face id
  // faces
  std::vector<int> fb0, fb1, fb2;
  std::vector<Vec> farea, fnorm;
  std::vector<double> fa;
  // subunits
  std::vector<int> centre;  // global id of each subunit's centre bead
  int centreLocal = 0;
  std::vector<std::vector<int>> lj_a;  // rows 1 and 2 used
  std::vector<Thermo> bath;
  double particle_ke = 0;
  std::default_random_engine gen;
  std::uniform_real_distribution<> distr;
  bool list_overflow = false;
  long long list_entries = 0;
  // encapsulation big beads (src/md.cpp:183-185): immobile; one diameter, charge and
  // hard-core distance for all of them
  std::vector<Vec> bpos;
  double bsig = 0, bq = 0, bshc = 0;
  std::vector<double> bce, bne;  // energies of each big bead

  Sys() : gen(69521u), distr(-0.5, 0.5) {}

  Vec dist(int a, int b) const {  // src/functions.h:42-56
    Vec r(pos[a].x - pos[b].x, pos[a].y - pos[b].y, pos[a].z - pos[b].z);
    min_image(r);
    return r;
  }
  void min_image(Vec& r) const {
    if (r.x > hbox.x) r.x -= box.x; else if (r.x < -hbox.x) r.x += box.x;
    if (r.y > hbox.y) r.y -= box.y; else if (r.y < -hbox.y) r.y += box.y;
    if (r.z > hbox.z) r.z -= box.z; else if (r.z < -hbox.z) r.z += box.z;
  }

  bool init(const char* path, const Cli& c) {
    cli = c;
    if (!read_template(path, t)) return false;
    Ns = c.n_subunits;
    N = Ns * t.np;
    // src/md.cpp:135-173
    const double Avagadro = 6.022e23, Boltzmann = 1.3806e-23, Pi = 3.14159;
    double SIenergy = c.K * Boltzmann;
    SItime = std::sqrt(t.SIsigma * t.SIsigma * t.SImass / SIenergy);
    T = c.K / double(298);
    box_x = std::pow((double(Ns) * 1000 / (c.C_uM * std::pow(t.SIsigma, 3) * Avagadro)), 1.0 / 3.0);
    ecut = 2.5 * (t.SIsigma / 1e-9);
    lb = (0.701e-9) / t.SIsigma;
    ni = c.c_mM * Avagadro * t.SIsigma * t.SIsigma * t.SIsigma;
    kappa = std::sqrt(8 * Pi * ni * lb * qs * qs);
    double screen = 1 / kappa;
    ecut_el = screen * c.ecut_c;
    if (ecut_el > ecut) NListCutoff = ecut_el + c.L_skin; else NListCutoff = ecut + c.L_skin;
    NLV = (unsigned)std::ceil((NListCutoff + 0.5) * (NListCutoff + 0.5) * (NListCutoff + 0.5));
    if (NLV > (unsigned)N) NLV = (unsigned)N;
    // generate_lattice's box is a long double built from the same double expression
    long double bx = std::pow((double(Ns) * 1000 / (c.C_uM * std::pow(t.SIsigma, 3) * 6.022e23)), 1.0 / 3.0);
    box = Vec((R)bx, (R)bx, (R)bx);
    hbox = box.scale((R)0.5);
    // beads (src/initialize.cpp:220-231)
    pos.assign(N, Vec()); vel = pos; tforce = pos; sforce = pos; bforce = pos; fdrag = pos; fran = pos;
    m.resize(N); sigma.resize(N); q.resize(N); type.resize(N); unit.resize(N);
    se.assign(N, 0); ke.assign(N, 0); be.assign(N, 0); ne.assign(N, 0); ce.assign(N, 0);
    for (int s = 0; s < Ns; s++)
      for (int l = 0; l < t.np; l++) {
        int g = s * t.np + l;
        q[g] = (double)t.q[l]; type[g] = t.type[l]; sigma[g] = (double)t.sigma[l]; m[g] = (double)t.mass[l];
        unit[g] = s;
      }
    // edges (src/initialize.cpp:317-339)
    itsE.assign(N, std::vector<int>());
    int E = Ns * t.ne;
    eb0.resize(E); eb1.resize(E); etype.resize(E); elen0.resize(E); elength.assign(E, 0); elenvec.assign(E, Vec());
    eF.assign(E, std::vector<int>());
    for (int s = 0; s < Ns; s++)
      for (int j = 0; j < t.ne; j++) {
        int idx = s * t.ne + j;
        int g1 = t.e1[j] + t.np * s, g2 = t.e2[j] + t.np * s;
        eb0[idx] = g1; eb1[idx] = g2; etype[idx] = t.etype[j]; elen0[idx] = (double)t.len0[j];
        itsE[g1].push_back(idx); itsE[g2].push_back(idx);
      }
    // faces and the face<->edge association (src/initialize.cpp:348-404), per template
    int F = Ns * t.nf;
    fb0.resize(F); fb1.resize(F); fb2.resize(F); farea.assign(F, Vec()); fnorm.assign(F, Vec()); fa.assign(F, 0);
    std::vector<std::vector<int>> tEF(t.ne);
    for (int f = 0; f < t.nf; f++)
      for (int e = 0; e < t.ne; e++) {
        int a = t.t1[f], b = t.t2[f], cc = t.t3[f], p = t.e1[e], r = t.e2[e];
        if ((a == p && b == r) || (a == r && b == p)) tEF[e].push_back(f);
        if ((a == p && cc == r) || (a == r && cc == p)) tEF[e].push_back(f);
        if ((b == r && cc == p) || (b == p && cc == r)) tEF[e].push_back(f);
      }
    for (int s = 0; s < Ns; s++) {
      for (int f = 0; f < t.nf; f++) {
        int idx = s * t.nf + f;
        fb0[idx] = t.t1[f] + t.np * s; fb1[idx] = t.t2[f] + t.np * s; fb2[idx] = t.t3[f] + t.np * s;
      }
      for (int e = 0; e < t.ne; e++)
        for (int f : tEF[e]) eF[s * t.ne + e].push_back(s * t.nf + f);
    }
    for (int e = 0; e < t.ne; e++)
      if (t.etype[e] != 0 && tEF[e].size() != 2) {
        g_err = "bending edge without exactly two faces"; return false;
      }
    lj_a.assign(3, std::vector<int>());
    for (int k = 0; k < t.na; k++) { lj_a[0].push_back(k); lj_a[1].push_back(t.la1[k]); lj_a[2].push_back(t.la2[k]); }
    itsN.assign(N, std::vector<int>(NLV, -1));
    centre.assign(Ns, 0);
    // NH chain (src/md.cpp:187-196)
    bath.clear();
    if (!c.brownian) {
      if (c.q_chain == 1) bath.push_back(Thermo(0, T, 3 * (unsigned long)N));
      else {
        bath.push_back(Thermo(c.Q, T, 3 * (unsigned long)N));
        while ((int)bath.size() != c.q_chain - 1) bath.push_back(Thermo(1, T, 1));
        bath.push_back(Thermo(0, T, 3 * (unsigned long)N));
      }
    }
    return true;
  }

  // centre bead from subunit 0's current positions (src/initialize.cpp:266-288)
  void find_centre() {
    Vec cen;
    for (int i = 0; i < t.np; i++) cen += pos[i];
    cen = Vec(cen.x / t.np, cen.y / t.np, cen.z / t.np);
    long double oldD = 1000;
    int id = 0;
    for (int i = 0; i < t.np; i++) {
      Vec r = pos[i] - cen;
      Vec hb(box.x / 2, box.y / 2, box.z / 2);
      if (r.x > hb.x) r.x -= box.x; else if (r.x < -hb.x) r.x += box.x;
      if (r.y > hb.y) r.y -= box.y; else if (r.y < -hb.y) r.y += box.y;
      if (r.z > hb.z) r.z -= box.z; else if (r.z < -hb.z) r.z += box.z;
      long double d = r.mag2();
      if (d < oldD) { oldD = d; id = i; }
    }
    centreLocal = id;
    for (int s = 0; s < Ns; s++) centre[s] = s * t.np + id;
  }

  void set_state(const double* p, const double* v) {
    for (int i = 0; i < N; i++) {
      pos[i] = Vec(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
      vel[i] = v ? Vec(v[3 * i], v[3 * i + 1], v[3 * i + 2]) : Vec();
    }
    find_centre();
    for (auto& b : bath) { b.xi = 0; b.eta = 0; }
    gen.seed(69521u);
    distr.reset();
  }

  // ---------------------------------------------------------------- bonded geometry
  void update_length(int e) {  // src/edge.cpp:35-49
    Vec lv = pos[eb0[e]] - pos[eb1[e]];
    min_image(lv);
    elenvec[e] = lv;
    elength[e] = (double)lv.mag();
  }
  void update_area_normal(int f) {  // src/face.cpp:11-17
    Vec first = dist(fb1[f], fb0[f]);
    Vec second = dist(fb2[f], fb1[f]);
    farea[f] = first.cross(second).scale((R)0.5);
    fa[f] = (double)farea[f].mag();
    fnorm[f] = farea[f].scale((R)(1 / fa[f]));
  }
  int across(int f, int e) const {  // src/face.cpp:27-34
    int b[3] = {fb0[f], fb1[f], fb2[f]};
    for (int i = 0; i < 3; i++)
      if (b[i] != eb0[e] && b[i] != eb1[e]) return b[i];
    return -1;
  }
  void stretching_force(int b, double ks) {  // src/bead.cpp:17-23
    Vec sf;
    for (int e : itsE[b]) {
      int other = (eb0[e] == b) ? eb1[e] : eb0[e];
      Vec r = dist(b, other);
      sf += r.scale((R)((elen0[e] - elength[e]) * ks / elength[e]));
    }
    sforce[b] = sf;
  }
  void stretching_energy(int b, double ks) {  // src/bead.cpp:9-14
    double s = 0;
    for (int e : itsE[b]) s += (0.25 * ks * (elen0[e] - elength[e]) * (elen0[e] - elength[e]));
    se[b] = s;
  }
  void flip_if_negative(int e) {  // src/edge.cpp:59-62, 78-81
    int F0 = eF[e][0], F1 = eF[e][1];
    if (fnorm[F0].dot(fnorm[F1]) < 0) {
      fnorm[F1] = fnorm[F1].scale((R)-1);
      farea[F1] = farea[F1].scale((R)-1);
    }
  }
  Vec gradS(int e, int wrt, const Vec& r13, const Vec& r01, const Vec& r12, const Vec& r02,
            const Vec& r03) const {  // src/edge.cpp:98-161
    Vec res;
    auto formula = [](const Vec& d23, const Vec& d32, const Vec& d20, const Vec& d21) {
      Vec a = d20.cross(d21), b = d20.cross(d32), out;
      out.x = d23.y * a.z - d21.y * b.z + d32.z * a.y + d21.z * b.y;
      out.y = d32.x * a.z + d21.x * b.z + d23.z * a.x - d21.z * b.x;
      out.z = d23.x * a.y - d21.x * b.y + d32.y * a.x + d21.y * b.x;
      return out;
    };
    if (wrt == eb0[e]) {
      res = formula(r13, r13.scale((R)-1), r01.scale((R)-1), r12);
    } else if (wrt == eb1[e]) {
      res = formula(r03, r03.scale((R)-1), r01, r02);
    } else {
      Vec r1x;
      int F0 = eF[e][0], F1 = eF[e][1];
      int b0[3] = {fb0[F0], fb1[F0], fb2[F0]}, b1[3] = {fb0[F1], fb1[F1], fb2[F1]};
      for (int i = 0; i < 3; i++)
        if (b0[i] != eb0[e] && b0[i] != eb1[e] && b0[i] != wrt) r1x = r12.scale((R)-1);
      for (int i = 0; i < 3; i++)
        if (b1[i] != eb0[e] && b1[i] != eb1[e] && b1[i] != wrt) r1x = r13.scale((R)-1);
      res = r01.cross(r01.cross(r1x));
    }
    return res;
  }
  Vec grad0(int e, int wrt, const Vec& r01, const Vec& r12, const Vec& r02) const {  // src/edge.cpp:163-174
    Vec res;
    int F0 = eF[e][0], F1 = eF[e][1];
    if (wrt == eb0[e]) res = r12.cross(r02.cross(r01)).scale((R)-1);
    else if (wrt == eb1[e]) res = r02.cross(r12.cross(r01));
    else if (wrt == across(F0, e)) res = r01.cross(r02.cross(r01)).scale((R)-1);
    else if (wrt == across(F1, e)) res = Vec();
    return res.scale((R)(0.25 / fa[F0]));
  }
  Vec grad1(int e, int wrt, const Vec& r13, const Vec& r01, const Vec& r03) const {  // src/edge.cpp:177-188
    Vec res;
    int F0 = eF[e][0], F1 = eF[e][1];
    if (wrt == eb0[e]) res = r13.cross(r03.cross(r01)).scale((R)-1);
    else if (wrt == eb1[e]) res = r03.cross(r13.cross(r01));
    else if (wrt == across(F1, e)) res = r01.cross(r03.cross(r01)).scale((R)-1);
    else if (wrt == across(F0, e)) res = Vec();
    return res.scale((R)(0.25 / fa[F1]));
  }
  void bending_forces(int e, double Kb) {  // src/edge.cpp:69-96
    int F0 = eF[e][0], F1 = eF[e][1];
    int r[4] = {eb0[e], eb1[e], across(F0, e), across(F1, e)};
    flip_if_negative(e);
    double S = (double)(farea[F0].dot(farea[F1]) * 4);
    Vec d01 = elenvec[e], d02 = dist(r[0], r[2]), d03 = dist(r[0], r[3]);
    Vec d12 = dist(r[1], r[2]), d13 = dist(r[1], r[3]);
    for (int i = 0; i < 4; i++) {
      Vec g = gradS(e, r[i], d13, d01, d12, d02, d03) -
              (grad0(e, r[i], d01, d12, d02).scale((R)(S / fa[F0])) +
               grad1(e, r[i], d13, d01, d03).scale((R)(S / fa[F1])));
      bforce[r[i]] += g.scale((R)(Kb / (fa[F0] * fa[F1] * 4)));
    }
  }
  void bending_energy(int e, double Kb) {  // src/edge.cpp:51-67
    int F0 = eF[e][0], F1 = eF[e][1];
    int r[4] = {eb0[e], eb1[e], across(F0, e), across(F1, e)};
    flip_if_negative(e);
    for (int i = 0; i < 4; i++) be[r[i]] += Kb * 0.25 * (double)(1 - (fnorm[F0].dot(fnorm[F1])));
  }

  // ---------------------------------------------------------------- neighbour list
  void update_pairlist(int i) {  // src/functions.cpp:69-117
    Vec rv;
    long double r2 = 0.0;
    double NL2 = NListCutoff * NListCutoff;
    int np = t.np;
    std::vector<unsigned> NLindex(np, 0);
    for (int ii = 0; ii < np; ii++) std::fill(itsN[i * np + ii].begin(), itsN[i * np + ii].end(), -1);
    for (int j = 0; j < Ns; j++) {
      if (i != j) {
        rv = pos[centre[i]] - pos[centre[j]];
        min_image(rv);
        r2 = rv.mag2();
      } else r2 = 0;
      if (r2 < (NL2 * 2.25)) {
        for (int ii = 0; ii < np; ii++) {
          int gi = i * np + ii;
          for (int jj = 0; jj < np; jj++) {
            int gj = j * np + jj;
            bool elec = (gi != gj && q[gi] != 0 && q[gj] != 0);
            bool lj = i != j;
            if (elec || lj) {
              // the reference compares the local index with the global centre id
              if (ii != centre[i] && jj != centre[j]) {
                rv = pos[gi] - pos[gj];
                min_image(rv);
                r2 = rv.mag2();
              }
              if (r2 < NL2) {
                if (NLindex[ii] < NLV) itsN[gi][NLindex[ii]] = gj;
                else list_overflow = true;
                NLindex[ii] += 1;
              }
            }
          }
        }
      }
    }
  }

  // ---------------------------------------------------------------- forces
  // src/forces.cpp:13-171. Optional per-term outputs (stretch, bend, Yukawa, LJ).
  void forces(bool rebuild, std::vector<Vec>* Fe_out = nullptr, std::vector<Vec>* Flj_out = nullptr) {
    const int np = t.np;
    const double ecut_el2 = ecut_el * ecut_el, ecut2 = ecut * ecut, replj2 = 1.12246205 * 1.12246205;
    const double ks = cli.ks, kb = cli.kb, elj_att = cli.elj_att;
    std::vector<Vec> forvec(N);
    long long entries = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : entries)
    for (int i = 0; i < Ns; i++) {
      Vec r_vec;
      long double r2 = 0.0;
      for (int e = i * t.ne; e < (i + 1) * t.ne; e++) update_length(e);
      for (int f = i * t.nf; f < (i + 1) * t.nf; f++) update_area_normal(f);
      for (int b = i * np; b < (i + 1) * np; b++) { stretching_force(b, ks); bforce[b] = Vec(); }
      for (int e = i * t.ne; e < (i + 1) * t.ne; e++)
        if (etype[e] != 0) bending_forces(e, kb);
      if (rebuild) update_pairlist(i);
      for (int ii = 0; ii < np; ii++) {
        int ib = i * np + ii;
        Vec eF_, ljF;
        for (unsigned jj = 0; jj < itsN[ib].size(); jj++) {
          int jb = itsN[ib][jj];
          if (jb == -1) break;
          entries++;
          bool electrostatic = (ib != jb && q[ib] != 0 && q[jb] != 0);
          bool lj = (i != unit[jb]);
          R r = 0.0;
          if (electrostatic || lj) {
            r_vec = pos[ib] - pos[jb];
            min_image(r_vec);
            r2 = r_vec.mag2();
          }
          double sig1 = sigma[ib], sig2 = sigma[jb];
          double shc = (sig1 + sig2) / 2;
          double yukawa_scale = std::exp(kappa * shc) / ((1 + kappa * sig1 / 2) * (1 + kappa * sig2 / 2));
          if ((electrostatic && r2 < ecut_el2 && ecut_el != 0) || (electrostatic && ecut_el == 0)) {
            r = std::sqrt((R)r2);
            eF_ += r_vec.scale((R)((q[ib] * q[jb] * lb * std::exp(-kappa * r) / ((R)r2)) * (kappa + 1 / r) *
                                   yukawa_scale));
          }
          if (r2 >= ecut2 && r2 >= (replj2 * shc * shc)) continue;
          if (lj) {
            bool att = false;
            for (size_t k = 0; k < lj_a[0].size(); k++)
              if (type[ib] == lj_a[1][k] && type[jb] == lj_a[2][k]) { att = true; break; }
            if (att && r2 < ecut2) {
              double test = (type[jb] == 7) ? 50 : elj_att;
              double r6 = (double)(r2 * r2 * r2);
              double sigma6 = sig1 * sig1 * sig1 * sig1 * sig1 * sig1;
              ljF += r_vec.scale((R)(48 * test * ((sigma6 / r6) * ((sigma6 / r6) - 0.5)) * (1 / ((R)r2))));
            } else if (!att && r2 < (replj2 * shc * shc)) {
              double r6 = (double)(r2 * r2 * r2);
              double sigma6 = shc * shc * shc * shc * shc * shc;
              double elj = 1;
              ljF += r_vec.scale((R)(48 * elj * ((sigma6 / r6) * ((sigma6 / r6) - 0.5)) * (1 / ((R)r2))));
            }
          }
        }
        Vec eljf = eF_ + ljF;
        forvec[ib] = sforce[ib] + bforce[ib] + eljf;
        if (Fe_out) (*Fe_out)[ib] = eF_;
        if (Flj_out) (*Flj_out)[ib] = ljF;
      }
    }
    list_entries = entries;
    for (int i = 0; i < N; i++) tforce[i] = forvec[i];
  }

  // src/forces.cpp:174-236: big bead i on subunit bead j, i outer and j inner. Only the
  // big-bead terms go into the bead's eforce / ljforce there, and its total becomes
  // tforce + ljforce + eforce.
  void forces_bigbead() {
    const int nbig = (int)bpos.size();
    if (nbig == 0) return;
    std::vector<Vec> ef(N), lf(N);
    const double shc = bshc, shc6 = std::pow(shc, 6);
    for (int i = 0; i < nbig; i++)
      for (int j = 0; j < N; j++) {
        const bool electrostatic = (q[j] != 0);
        Vec rv = bpos[i] - pos[j];
        min_image(rv);
        const R r = rv.mag(), r2 = rv.mag2();
        const double sig1 = bsig, sig2 = sigma[j], sigavg = (sig1 + sig2) / 2, delta = sigavg - shc;
        const double r6 = (double)std::pow((R)(r - delta), 6);
        const double ys = std::exp(kappa * sigavg) / ((1 + kappa * sig1 / 2) * (1 + kappa * sig2 / 2));
        Vec eF, ljF;
        if (electrostatic && r2 < (ecut_el * ecut_el))
          eF = rv.scale((R)((bq * q[j] * lb * std::exp(-kappa * r) / (r2)) * (kappa + 1 / r) * ys));
        ef[j] -= eF;
        if (r2 < (delta + 1.12246205 * shc) * (delta + 1.12246205 * shc))
          ljF = rv.scale((R)(48 * ((shc6 / r6) * ((shc6 / r6) - 0.5)) * (1 / r) * (1 / (r - delta))));
        lf[j] -= ljF;
      }
    for (int j = 0; j < N; j++) tforce[j] = tforce[j] + lf[j] + ef[j];
  }
  void all_forces(bool rebuild, std::vector<Vec>* Fe_out = nullptr, std::vector<Vec>* Flj_out = nullptr) {
    forces(rebuild, Fe_out, Flj_out);
    forces_bigbead();
  }

  // ---------------------------------------------------------------- energies
  long double particle_kinetic_energy() {  // src/energies.cpp:13-22
    for (int i = 0; i < N; i++) { R v = vel[i].mag(); ke[i] = (double)(0.5 * m[i] * v * v); }
    long double k = 0.0;
    for (int i = 0; i < N; i++) k += ke[i];
    return k;
  }
  void lj_es_energies() {  // src/energies.cpp:24-93, all i<j pairs
    const double ecut3 = ecut * ecut * ecut, ecut6 = ecut3 * ecut3, elj_att = cli.elj_att;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < N; i++) {
      double cei = ce[i], nei = ne[i];
      for (int j = i + 1; j < N; j++) {
        bool electrostatic = (i != j && q[i] != 0 && q[j] != 0);
        bool lj = unit[i] != unit[j];
        R r = 0.0;
        if (electrostatic || lj) {
          Vec rv = pos[i] - pos[j];
          min_image(rv);
          r = rv.mag();
        }
        double sig1 = sigma[i], sig2 = sigma[j];
        double shc = (sig1 + sig2) / 2;
        double ys = std::exp(kappa * shc) / ((1 + kappa * sig1 / 2) * (1 + kappa * sig2 / 2));
        if (electrostatic && r < ecut_el && ecut_el != 0) {
          cei += (double)(ys * ((q[i] * q[j] * lb * std::exp(-kappa * r)) / (r) -
                               ((q[i] * q[j] * lb * std::exp(-kappa * ecut_el)) / (ecut_el))));
        } else if (electrostatic && ecut_el == 0) {
          cei += (double)((q[i] * q[j] * lb * std::exp(-kappa * r) * ys) / (r));
        }
        double sigma3 = shc * shc * shc, sigma6 = sigma3 * sigma3;
        double r3 = (double)(r * r * r), r6 = r3 * r3;
        if (r >= ecut && r >= (1.12246205 * shc)) continue;
        if (lj) {
          bool att = false;
          for (size_t k = 0; k < lj_a[0].size(); k++)
            if (type[i] == lj_a[1][k] && type[j] == lj_a[2][k]) { att = true; break; }
          if (r < (1.12246205 * shc) && !att) {
            double elj = 1;
            nei += ((4 * elj * (sigma6 / r6) * ((sigma6 / r6) - 1)) + elj);
          } else if (r < ecut && att) {
            nei += (((4 * elj_att * (sigma6 / r6) * ((sigma6 / r6) - 1)) -
                     (4 * elj_att * ((sigma6 / ecut6) * ((sigma6 / ecut6) - 1)))));
          }
        }
      }
      ce[i] = cei;
      ne[i] = nei;
    }
  }
  // src/energies.cpp:96-136: each big-bead / subunit-bead term is added to both beads'
  // energies (the subunit bead's at full value)
  void energies_bigbead() {
    const int nbig = (int)bpos.size();
    bce.assign(nbig, 0.0);
    bne.assign(nbig, 0.0);
    const double shc = bshc, shc6 = std::pow(shc, 6);
    for (int i = 0; i < nbig; i++)
      for (int j = 0; j < N; j++) {
        const bool electrostatic = (q[j] != 0);
        Vec rv = bpos[i] - pos[j];
        min_image(rv);
        const R r = rv.mag();
        const double sig1 = bsig, sig2 = sigma[j], sigavg = (sig1 + sig2) / 2, delta = sigavg - shc;
        const double r6 = (double)std::pow((R)(r - delta), 6);
        const double ys = std::exp(kappa * sigavg) / ((1 + kappa * sig1 / 2) * (1 + kappa * sig2 / 2));
        double cej = 0, nej = 0;
        if (electrostatic && r < ecut_el) cej = (double)((bq * q[j] * lb * std::exp(-kappa * r) * ys) / (r));
        bce[i] += cej;
        ce[j] += cej;
        if (r < delta + 1.12246205 * shc) nej = ((4 * (shc6 / r6) * ((shc6 / r6) - 1)) + 1);
        bne[i] += nej;
        ne[j] += nej;
      }
  }
  // src/md.cpp:553-623: returns ke, se, be, ne, ce, tpe, tke (totals)
  void energies(double* out) {
    for (size_t e = 0; e < eb0.size(); e++) update_length((int)e);  // dress_up
    for (size_t f = 0; f < fb0.size(); f++) update_area_normal((int)f);
    for (int i = 0; i < N; i++) { be[i] = 0; ne[i] = 0; ce[i] = 0; }
    for (int s = 0; s < Ns; s++) {
      for (int b = s * t.np; b < (s + 1) * t.np; b++) {
        stretching_energy(b, cli.ks);
        R v = vel[b].mag();
        ke[b] = (double)(0.5 * m[b] * v * v);
      }
      for (int e = s * t.ne; e < (s + 1) * t.ne; e++)
        if (etype[e] != 0) bending_energy(e, cli.kb);
    }
    lj_es_energies();
    energies_bigbead();
    double S = 0, K = 0, L = 0, B = 0, Ce = 0, tp = 0, tk = 0;
    for (int i = 0; i < N; i++) { S += se[i]; K += ke[i]; L += ne[i]; B += be[i]; Ce += ce[i]; }
    for (auto& b : bath) {
      b.pe = b.dof * b.kB * b.T * b.eta;
      b.ke = 0.5 * b.Q * b.xi * b.xi;
      tp += b.pe; tk += b.ke;
    }
    out[0] = K; out[1] = S; out[2] = B; out[3] = L; out[4] = Ce; out[5] = tp; out[6] = tk;
  }

  // ---------------------------------------------------------------- step loop body
  void therm_update_velocity(int i, double dt, double expfac) {  // src/bead.h:113-117
    vel[i] = vel[i].scale((R)expfac) + tforce[i].scale((R)(0.5 * dt * (std::sqrt(expfac)) / m[i]));
  }
  void update_position(int i, double dt) {  // src/bead.h:72-81
    Vec& p = pos[i];
    p = p + vel[i].scale((R)dt);
    if (p.x > hbox.x) p.x = p.x - box.x; else if (p.x < -hbox.x) p.x = p.x + box.x;
    if (p.y > hbox.y) p.y = p.y - box.y; else if (p.y < -hbox.y) p.y = p.y + box.y;
    if (p.z > hbox.z) p.z = p.z - box.z; else if (p.z < -hbox.z) p.z = p.z + box.z;
  }
  void update_velocity(int i, double dt) { vel[i] = vel[i] + tforce[i].scale((R)(0.5 * dt / m[i])); }

  // initial force call and kinetic energy (src/md.cpp:241, 258)
  void start() {
    list_overflow = false;
    all_forces(true);
    particle_ke = (double)particle_kinetic_energy();
  }

  void step(long long a) {  // src/md.cpp:400-514
    const double dt = cli.dt, damp = cli.damp;
    double expfac_real = 0;
    if (!cli.brownian) {
      for (int i = (int)bath.size() - 1; i > -1; i--) update_chain_xi(i, bath, dt, particle_ke);
      for (auto& b : bath) b.update_eta(dt);
      expfac_real = std::exp(-0.5 * dt * bath[0].xi);
      for (int i = 0; i < N; i++) { therm_update_velocity(i, dt, expfac_real); update_position(i, dt); }
    } else {
      for (int i = 0; i < N; i++) {
        fdrag[i] = vel[i].scale((R)(-1.0 * m[i] / damp));
        update_velocity(i, dt);
        update_position(i, dt);
      }
    }
    bool rebuild = (a % cli.B_build == 0);
    all_forces(rebuild);
    if (!cli.brownian) {
      for (int i = 0; i < N; i++) therm_update_velocity(i, dt, expfac_real);
      particle_ke = (double)particle_kinetic_energy();
      for (auto& b : bath) b.update_eta(dt);
      for (unsigned i = 0; i < bath.size(); i++) update_chain_xi(i, bath, dt, particle_ke);
    } else {
      for (int i = 0; i < N; i++) {
        double amp = std::sqrt((m[i] * 24.0) / (damp * dt));
        fran[i].x = amp * distr(gen);
        fran[i].y = amp * distr(gen);
        fran[i].z = amp * distr(gen);
      }
      Vec avg;
      for (int i = 0; i < N; i++) avg = avg + fran[i];
      avg = avg.scale((R)(1.0 / N));
      for (int i = 0; i < N; i++) fran[i] = fran[i] - avg;
      // the big beads' noise is drawn next and never used (src/md.cpp:482-491)
      for (size_t k = 0; k < 3 * bpos.size(); k++) distr(gen);
      for (int i = 0; i < N; i++) {
        tforce[i] = tforce[i] + fran[i] + fdrag[i];
        update_velocity(i, dt);
      }
    }
    Vec av;
    for (int i = 0; i < N; i++) av = av + vel[i];
    av = av.scale((R)(1.0 / N));
    for (int i = 0; i < N; i++) vel[i] = vel[i] - av;
    last_vcom = (double)av.mag();
  }
  double last_vcom = 0;

  // Oligomers of the current positions, as the mass-spectrum pass of src/analysis.cpp:143-186:
  // each subunit not yet in an oligomer seeds a new one, which grows round by round over the
  // members added in the previous round, adding every higher-numbered unassigned subunit that
  // has a bead within `contact` (dist(...).GetMagnitude() < 2 there) of the member's beads.
  // label[s] = the seed (smallest subunit) of s's oligomer; ms[k - 1] = oligomers of k subunits.
  int oligomers(double contact, int* label, int* ms) const {
    const int np = t.np;
    std::vector<int> owner(Ns, -1);
    std::vector<std::vector<int>> olig;
    for (int i = 0; i < Ns; i++) {
      if (owner[i] != -1) continue;
      const int idx = (int)olig.size();
      olig.push_back(std::vector<int>(1, i));
      owner[i] = idx;
      size_t oldsize = 0;
      while (oldsize < olig[idx].size()) {
        const size_t n = oldsize;
        oldsize = olig[idx].size();
        for (size_t j = n; j < oldsize; j++) {
          const int g = olig[idx][j];
          for (int k = i + 1; k < Ns; k++) {
            if (owner[k] != -1) continue;
            bool touch = false;
            for (int m = 0; m < np && !touch; m++)
              for (int nn = 0; nn < np; nn++)
                if (dist(g * np + m, k * np + nn).mag() < (R)contact) { touch = true; break; }
            if (touch) { olig[idx].push_back(k); owner[k] = idx; }
          }
        }
      }
    }
    for (int s = 0; s < Ns; s++) {
      if (label) label[s] = olig[owner[s]][0];
      if (ms) ms[s] = 0;
    }
    if (ms)
      for (auto& o : olig) ms[o.size() - 1] += 1;
    return (int)olig.size();
  }
};

// single-hinge literal bending force for a 4-bead system (tests): builds a one-subunit
// system whose template has one bending edge (0,1) and faces (0,1,2), (0,1,3).
}  // namespace

struct OrcHandle {
  int ld;  // 1 = long double mode
  SysBase* s;
};

template <typename F>
static auto dispatch(OrcHandle* h, F f) {
  if (h->ld) return f(*static_cast<Sys<long double>*>(h->s));
  return f(*static_cast<Sys<double>*>(h->s));
}

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Create a system from a template file and the reference's CLI values.
// cli_d = {C_uM, c_mM, ks, kb, elj_att, ecut_c, L_skin, dt, K, Q, damp}
// cli_i = {n_subunits, B_build, q_chain, brownian}
void* orc_create(const char* path, const double* cli_d, const int* cli_i, int long_double_mode) {
  Cli c;
  c.C_uM = cli_d[0]; c.c_mM = cli_d[1]; c.ks = cli_d[2]; c.kb = cli_d[3]; c.elj_att = cli_d[4];
  c.ecut_c = cli_d[5]; c.L_skin = cli_d[6]; c.dt = cli_d[7]; c.K = cli_d[8]; c.Q = cli_d[9]; c.damp = cli_d[10];
  c.n_subunits = cli_i[0]; c.B_build = cli_i[1]; c.q_chain = cli_i[2]; c.brownian = cli_i[3];
  OrcHandle* h = new OrcHandle;
  h->ld = long_double_mode;
  bool ok;
  if (long_double_mode) { auto* s = new Sys<long double>(); ok = s->init(path, c); h->s = s; }
  else { auto* s = new Sys<double>(); ok = s->init(path, c); h->s = s; }
  if (!ok) { delete h->s; delete h; return nullptr; }
  return h;
}

void orc_destroy(void* hv) {
  OrcHandle* h = (OrcHandle*)hv;
  if (!h) return;
  delete h->s;
  delete h;
}

// out: {box, ecut, lb, kappa, ecut_el, NListCutoff, T, SItime, NLV, N, Ns, np, centre_local}
void orc_params(void* hv, double* out) {
  dispatch((OrcHandle*)hv, [&](auto& s) {
    out[0] = (double)s.box.x; out[1] = s.ecut; out[2] = s.lb; out[3] = s.kappa; out[4] = s.ecut_el;
    out[5] = s.NListCutoff; out[6] = s.T; out[7] = s.SItime; out[8] = s.NLV; out[9] = s.N; out[10] = s.Ns;
    out[11] = s.t.np; out[12] = s.centreLocal;
    return 0;
  });
}

// Positions and velocities, AoS [N][3]. Resets thermostat and BD stream as a (re)start
// of the reference does, runs the initial force call and kinetic energy.
int orc_set_state(void* hv, const double* pos, const double* vel) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    s.set_state(pos, vel);
    s.start();
    return s.list_overflow ? -2 : 0;
  });
}

// Sets the chain (xi, eta for each of the n thermostats) and the particle kinetic energy that
// the next step's first chain update uses (src/md.cpp:400-403 reads the value measured at the
// end of the previous step), after orc_set_state: together they continue a run that another
// program stopped mid-trajectory. n must equal the chain length.
int orc_set_thermostat(void* hv, const double* xi, const double* eta, int n, double particle_ke) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    if (n != (int)s.bath.size()) return -1;
    for (int k = 0; k < n; k++) { s.bath[k].xi = xi[k]; s.bath[k].eta = eta[k]; }
    s.particle_ke = particle_ke;
    return 0;
  });
}

// Test knob: a neighbour-list capacity per bead other than the reference's fixed
// ceil((NListCutoff + 0.5)^3) (src/md.cpp:231), for states denser than that capacity was sized
// for (the reference would overflow it there); the list algorithm itself is unchanged. Call
// before orc_set_state.
int orc_set_list_capacity(void* hv, int nlv) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    if (nlv <= 0) return -1;
    s.NLV = (unsigned)std::min(nlv, s.N);
    s.itsN.assign(s.N, std::vector<int>(s.NLV, -1));
    return 0;
  });
}

int orc_get_state(void* hv, double* pos, double* vel, double* xi, double* eta) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    for (int i = 0; i < s.N; i++) {
      if (pos) { pos[3 * i] = (double)s.pos[i].x; pos[3 * i + 1] = (double)s.pos[i].y; pos[3 * i + 2] = (double)s.pos[i].z; }
      if (vel) { vel[3 * i] = (double)s.vel[i].x; vel[3 * i + 1] = (double)s.vel[i].y; vel[3 * i + 2] = (double)s.vel[i].z; }
    }
    for (size_t k = 0; k < s.bath.size(); k++) {
      if (xi) xi[k] = s.bath[k].xi;
      if (eta) eta[k] = s.bath[k].eta;
    }
    return (int)s.bath.size();
  });
}

// Forces at the current positions (a forceCalculation call). Outputs are [N][3]; any may
// be NULL: total, stretch, bend, Yukawa, LJ. rebuild != 0 rebuilds the list first.
// Does not disturb the integrator's tforce.
int orc_forces(void* hv, int rebuild, double* F, double* Fs, double* Fb, double* Fe, double* Flj) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    typedef typename std::remove_reference<decltype(s)>::type S;
    std::vector<typename S::Vec> keep = s.tforce, fe(s.N), flj(s.N);
    s.list_overflow = false;
    s.all_forces(rebuild != 0, &fe, &flj);
    for (int i = 0; i < s.N; i++) {
      auto put = [&](double* o, const typename S::Vec& v) {
        if (o) { o[3 * i] = (double)v.x; o[3 * i + 1] = (double)v.y; o[3 * i + 2] = (double)v.z; }
      };
      put(F, s.tforce[i]); put(Fs, s.sforce[i]); put(Fb, s.bforce[i]); put(Fe, fe[i]); put(Flj, flj[i]);
    }
    s.tforce = keep;
    return s.list_overflow ? -2 : 0;
  });
}

// Totals {ke, se, be, ne, ce, tpe, tke}; per_bead (nullable) [N][5] = se, be, ne, ce, ke.
int orc_energies(void* hv, double* out, double* per_bead) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    s.energies(out);
    if (per_bead)
      for (int i = 0; i < s.N; i++) {
        per_bead[5 * i] = s.se[i]; per_bead[5 * i + 1] = s.be[i]; per_bead[5 * i + 2] = s.ne[i];
        per_bead[5 * i + 3] = s.ce[i]; per_bead[5 * i + 4] = s.ke[i];
      }
    return 0;
  });
}

// Runs loop iterations a = a_start .. a_start + n_steps - 1 (no output). Returns 0, or -2
// if a list outgrew its capacity.
int orc_run(void* hv, long long a_start, int n_steps) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    s.list_overflow = false;
    for (int k = 0; k < n_steps; k++) s.step(a_start + k);
    return s.list_overflow ? -2 : 0;
  });
}

// Loop-only wall time (std::chrono) of n_steps iterations with n_threads OpenMP threads.
double orc_time_steps(void* hv, long long a_start, int n_steps, int n_threads) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  auto t0 = std::chrono::steady_clock::now();
  orc_run(hv, a_start, n_steps);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

long long orc_list_entries(void* hv) {
  return dispatch((OrcHandle*)hv, [&](auto& s) { return s.list_entries; });
}

// Neighbour list of bead i (entries before the -1 terminator); returns the count.
int orc_neighbours(void* hv, int i, int* out, int cap) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    int n = 0;
    for (unsigned k = 0; k < s.itsN[i].size(); k++) {
      if (s.itsN[i][k] == -1) break;
      if (n < cap) out[n] = s.itsN[i][k];
      n++;
    }
    return n;
  });
}

// Template-level hinge table: for each bending edge (template order) the 4 beads
// r0 = B0, r1 = B1, r2 = across(F0), r3 = across(F1), template-local ids. Returns count.
int orc_hinges(void* hv, int* out, int cap) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    int n = 0;
    for (int e = 0; e < s.t.ne; e++) {
      if (s.etype[e] == 0) continue;
      int F0 = s.eF[e][0], F1 = s.eF[e][1];
      if (n < cap) {
        out[4 * n] = s.eb0[e]; out[4 * n + 1] = s.eb1[e];
        out[4 * n + 2] = s.across(F0, e); out[4 * n + 3] = s.across(F1, e);
      }
      n++;
    }
    return n;
  });
}

// Oligomers at the current positions (contact distance 2 in the reference's analysis):
// label[Ns] = smallest subunit of each subunit's oligomer, ms[Ns] = oligomers by size
// (either may be NULL). Returns the number of oligomers.
int orc_oligomers(void* hv, double contact, int* label, int* ms) {
  return dispatch((OrcHandle*)hv, [&](auto& s) { return s.oligomers(contact, label, ms); });
}

// Encapsulation big beads: n immobile beads at xyz[3n] of diameter sigma, charge q (e) and
// hard-core distance shc; n = 0 removes them. Positions are taken through the minimum
// image only, as the reference does. Takes effect at the next force call.
int orc_set_big_beads(void* hv, int n, const double* xyz, double sigma, double q, double shc) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    typedef typename std::remove_reference<decltype(s)>::type S;
    s.bpos.clear();
    for (int i = 0; i < n; i++) s.bpos.push_back(typename S::Vec(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    s.bsig = sigma; s.bq = q; s.bshc = shc;
    return 0;
  });
}

// Per-bead diameter and charge (nb = N), either may be NULL.
int orc_bead_props(void* hv, double* sigma, double* q) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    for (int i = 0; i < s.N; i++) {
      if (sigma) sigma[i] = s.sigma[i];
      if (q) q[i] = s.q[i];
    }
    return s.N;
  });
}

// Energies of the big beads from the last orc_energies call: out = {sum of ce, sum of ne}.
int orc_big_energies(void* hv, double* out) {
  return dispatch((OrcHandle*)hv, [&](auto& s) {
    double c = 0, l = 0;
    for (size_t i = 0; i < s.bce.size(); i++) { c += s.bce[i]; l += s.bne[i]; }
    out[0] = c; out[1] = l;
    return 0;
  });
}

// generate_big_beads (src/initialize.cpp:566-680) without a restart file: srand(0), then per
// big bead random positions (three rand() calls, x, y, z) until one keeps a distance of at
// least (sigma_j + sigma) / 2 from every existing bead j and sigma from every big bead
// placed before (minimum image by rounding). Existing beads: pos[3 nb], sig[nb]; cubic
// box. The reference's positions and box are long double; so is the arithmetic here.
// Returns 0, or -1 (as the reference's exit) when a bead needs max_attempts tries.
int orc_place_big_beads(const double* pos, const double* sig, int nb, double box, int n, double sigma,
                        unsigned max_attempts, double* out) {
  typedef long double LD;
  const LD bx = box;
  const double R_squared = sigma * sigma;
  std::srand(0);
  int placed_n = 0;
  for (int i = 0; i < n; i++) {
    bool placed = false;
    unsigned attempts = 0;
    while (!placed && attempts < max_attempts) {
      const double x = (static_cast<double>(std::rand()) / RAND_MAX - 0.5) * bx;
      const double y = (static_cast<double>(std::rand()) / RAND_MAX - 0.5) * bx;
      const double z = (static_cast<double>(std::rand()) / RAND_MAX - 0.5) * bx;
      bool collision = false;
      for (int j = 0; j < nb && !collision; j++) {
        double dx = x - (LD)pos[3 * j];
        dx -= bx * std::floor(dx / bx + 0.5);
        double dy = y - (LD)pos[3 * j + 1];
        dy -= bx * std::floor(dy / bx + 0.5);
        double dz = z - (LD)pos[3 * j + 2];
        dz -= bx * std::floor(dz / bx + 0.5);
        const double dist_sq = dx * dx + dy * dy + dz * dz;
        const double min_dist = sig[j] / 2 + sigma / 2;
        if (dist_sq < min_dist * min_dist) collision = true;
      }
      for (int k = 0; k < placed_n && !collision; k++) {
        double dx = x - (LD)out[3 * k];
        dx -= bx * std::floor(dx / bx + 0.5);
        double dy = y - (LD)out[3 * k + 1];
        dy -= bx * std::floor(dy / bx + 0.5);
        double dz = z - (LD)out[3 * k + 2];
        dz -= bx * std::floor(dz / bx + 0.5);
        const double dist_sq = dx * dx + dy * dy + dz * dz;
        if (dist_sq < R_squared) collision = true;
      }
      if (!collision) {
        out[3 * placed_n] = x; out[3 * placed_n + 1] = y; out[3 * placed_n + 2] = z;
        placed_n++;
        placed = true;
      }
      attempts++;
    }
    if (attempts >= max_attempts) return -1;
  }
  return 0;
}

// The BD noise stream of the reference: n uniform(-0.5, 0.5) draws of
// std::default_random_engine(69521) after skipping `skip` draws of the distribution.
void orc_bd_uniforms(long long skip, int n, double* out) {
  std::default_random_engine gen(69521u);
  std::uniform_real_distribution<> distr(-0.5, 0.5);
  for (long long k = 0; k < skip; k++) distr(gen);
  for (int k = 0; k < n; k++) out[k] = distr(gen);
}

}  // extern "C"
