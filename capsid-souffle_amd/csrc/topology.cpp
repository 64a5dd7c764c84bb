// Host side: template reader, per-template topology, parameter derivation and the seeded
// initial-state generator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <unordered_map>

#include "sf_internal.hpp"

namespace sf {

thread_local std::string g_error;
void set_error(const std::string& s) { g_error = s; }

uint32_t lcg_pow(uint64_t e) {
  uint64_t result = 1, base = kLcgMul;
  while (e) {
    if (e & 1) result = (result * base) % kLcgMod;
    base = (base * base) % kLcgMod;
    e >>= 1;
  }
  return (uint32_t)result;
}

// Derived per-template tables. The bending hinge of an edge uses the two faces holding both
// of its beads, F0 the lower face index, and the across bead of a face is its first vertex
// not on the edge: the association src/initialize.cpp:348-404 builds and FACE::across
// (src/face.cpp:27-34) reads.
bool derive_template(sf_template& t) {
  const int np = t.np;
  if (np <= 0 || np > kBlock) { set_error("template must have 1..128 beads"); return false; }
  for (int l = 0; l < np; l++)
    if (t.type[l] < 0 || t.type[l] >= kMaxTypes) { set_error("bead type out of range 0..63"); return false; }
  for (int e = 0; e < t.ne; e++)
    if (t.e1[e] < 0 || t.e1[e] >= np || t.e2[e] < 0 || t.e2[e] >= np || t.e1[e] == t.e2[e]) {
      set_error("edge references an invalid bead"); return false;
    }
  for (int f = 0; f < t.nf; f++)
    if (t.f1[f] < 0 || t.f1[f] >= np || t.f2[f] < 0 || t.f2[f] >= np || t.f3[f] < 0 || t.f3[f] >= np) {
      set_error("face references an invalid bead"); return false;
    }
  // hinges
  t.hinges.clear();
  for (int e = 0; e < t.ne; e++) {
    if (t.etype[e] == 0) continue;
    int a = t.e1[e], b = t.e2[e];
    int faces[2], nfc = 0;
    for (int f = 0; f < t.nf; f++) {
      int v[3] = {t.f1[f], t.f2[f], t.f3[f]};
      bool ha = (v[0] == a || v[1] == a || v[2] == a), hb = (v[0] == b || v[1] == b || v[2] == b);
      if (ha && hb) {
        if (nfc < 2) faces[nfc] = f;
        nfc++;
      }
    }
    if (nfc != 2) { set_error("bending edge without exactly two faces"); return false; }
    int across[2];
    for (int k = 0; k < 2; k++) {
      int f = faces[k];
      int v[3] = {t.f1[f], t.f2[f], t.f3[f]};
      across[k] = -1;
      for (int i = 0; i < 3; i++)
        if (v[i] != a && v[i] != b) { across[k] = v[i]; break; }
      if (across[k] < 0) { set_error("degenerate face on a bending edge"); return false; }
    }
    t.hinges.push_back(a); t.hinges.push_back(b);
    t.hinges.push_back(across[0]); t.hinges.push_back(across[1]);
  }
  // per-bead edge incidence
  t.e_off.assign(np + 1, 0); t.e_other.clear(); t.e_len0.clear();
  for (int l = 0; l < np; l++) {
    for (int e = 0; e < t.ne; e++) {
      if (t.e1[e] == l) { t.e_other.push_back(t.e2[e]); t.e_len0.push_back(t.len0[e]); }
      else if (t.e2[e] == l) { t.e_other.push_back(t.e1[e]); t.e_len0.push_back(t.len0[e]); }
    }
    t.e_off[l + 1] = (int)t.e_other.size();
  }
  // per-bead hinge incidence
  const int nh = t.nh();
  t.h_off.assign(np + 1, 0); t.h_slot.clear();
  for (int l = 0; l < np; l++) {
    for (int h = 0; h < nh; h++)
      for (int k = 0; k < 4; k++)
        if (t.hinges[4 * h + k] == l) t.h_slot.push_back(4 * h + k);
    t.h_off[l + 1] = (int)t.h_slot.size();
  }
  // charged beads
  t.qslot.assign(np, -1); t.qlocal.clear();
  for (int l = 0; l < np; l++)
    if (t.q[l] != 0) { t.qslot[l] = (int)t.qlocal.size(); t.qlocal.push_back(l); }
  // attractive LJ table (ordered rows of lj_a, src/forces.cpp:119-125)
  t.att_mask.assign(kMaxTypes, 0ull);
  for (size_t k = 0; k < t.la1.size(); k++) {
    int a = t.la1[k], b = t.la2[k];
    if (a >= 0 && a < kMaxTypes && b >= 0 && b < kMaxTypes) t.att_mask[a] |= (1ull << b);
  }
  // centre bead and radii
  double cx = 0, cy = 0, cz = 0;
  for (int l = 0; l < np; l++) { cx += t.x[l]; cy += t.y[l]; cz += t.z[l]; }
  cx /= np; cy /= np; cz /= np;
  double best = 1e300;
  for (int l = 0; l < np; l++) {
    double d = (t.x[l] - cx) * (t.x[l] - cx) + (t.y[l] - cy) * (t.y[l] - cy) + (t.z[l] - cz) * (t.z[l] - cz);
    if (d < best) { best = d; t.centre = l; }
  }
  t.r_all = 0; t.r_q = 0;
  for (int l = 0; l < np; l++) {
    double d = std::sqrt((t.x[l] - t.x[t.centre]) * (t.x[l] - t.x[t.centre]) +
                         (t.y[l] - t.y[t.centre]) * (t.y[l] - t.y[t.centre]) +
                         (t.z[l] - t.z[t.centre]) * (t.z[l] - t.z[t.centre]));
    t.r_all = std::max(t.r_all, d);
    if (t.q[l] != 0) t.r_q = std::max(t.r_q, d);
  }
  t.diam = 0;
  for (int a = 0; a < np; a++)
    for (int b = a + 1; b < np; b++)
      t.diam = std::max(t.diam, std::sqrt((t.x[a] - t.x[b]) * (t.x[a] - t.x[b]) + (t.y[a] - t.y[b]) * (t.y[a] - t.y[b]) +
                                          (t.z[a] - t.z[b]) * (t.z[a] - t.z[b])));
  // body frame: fp1 = bead farthest from the centre bead, fp2 = bead farthest from that axis
  // (first index on ties)
  const int c = t.centre;
  double e1[3] = {0, 0, 0}, b1 = -1;
  for (int l = 0; l < np; l++) {
    const double u[3] = {t.x[l] - t.x[c], t.y[l] - t.y[c], t.z[l] - t.z[c]};
    const double n2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (n2 > b1 + 1e-12) { b1 = n2; t.fp1 = l; }
  }
  {
    const double u[3] = {t.x[t.fp1] - t.x[c], t.y[t.fp1] - t.y[c], t.z[t.fp1] - t.z[c]};
    const double n = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int k = 0; k < 3; k++) e1[k] = n > 0 ? u[k] / n : 0.0;
  }
  double b2 = -1;
  for (int l = 0; l < np; l++) {
    const double u[3] = {t.x[l] - t.x[c], t.y[l] - t.y[c], t.z[l] - t.z[c]};
    const double a = u[0] * e1[0] + u[1] * e1[1] + u[2] * e1[2];
    const double w2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2] - a * a;
    if (w2 > b2 + 1e-12) { b2 = w2; t.fp2 = l; }
  }
  double R[9];
  const double u1[3] = {t.x[t.fp1] - t.x[c], t.y[t.fp1] - t.y[c], t.z[t.fp1] - t.z[c]};
  const double u2[3] = {t.x[t.fp2] - t.x[c], t.y[t.fp2] - t.y[c], t.z[t.fp2] - t.z[c]};
  body_frame(u1, u2, R);
  t.qbody.assign(3 * t.qlocal.size(), 0.0);
  for (size_t k = 0; k < t.qlocal.size(); k++) {
    const int l = t.qlocal[k];
    const double u[3] = {t.x[l] - t.x[c], t.y[l] - t.y[c], t.z[l] - t.z[c]};
    for (int a = 0; a < 3; a++) t.qbody[3 * k + a] = R[a] * u[0] + R[3 + a] * u[1] + R[6 + a] * u[2];  // R^T u
  }
  t.frame_ok = b2 > 1e-12;
  return true;
}

}  // namespace sf

using namespace sf;

extern "C" {

const char* sf_last_error(void) { return g_error.c_str(); }

// Token reader of the reference's template format (src/initialize.cpp:46-128).
sf_template* sf_template_load(const char* path) {
  if (!path) { set_error("null path"); return nullptr; }
  std::ifstream in(path);
  if (!in) { set_error(std::string("cannot open template ") + path); return nullptr; }
  sf_template* t = new sf_template();
  std::string d;
  in >> d >> d >> t->bondlength >> d >> d >> t->si_mass >> d >> d >> t->si_sigma;
  in >> d >> d >> d >> d >> d >> t->np;
  for (int i = 0; i < 8; i++) in >> d;
  for (int i = 0; i < t->np && in; i++) {
    long long name, type;
    double x, y, z, c, r, m;
    in >> name >> x >> y >> z >> type >> c >> r >> m;
    t->x.push_back(x); t->y.push_back(y); t->z.push_back(z); t->type.push_back((int)type);
    t->q.push_back(c); t->sigma.push_back(r); t->mass.push_back(m);
  }
  in >> d >> d >> d >> d >> d >> t->ne;
  for (int i = 0; i < 5; i++) in >> d;
  for (int i = 0; i < t->ne && in; i++) {
    long long name, type; int a, b; double len;
    in >> name >> a >> b >> type >> len;
    t->e1.push_back(a); t->e2.push_back(b); t->etype.push_back((int)type); t->len0.push_back(len);
  }
  in >> d >> d >> d >> d >> d >> t->nf;
  for (int i = 0; i < 6; i++) in >> d;
  for (int i = 0; i < t->nf && in; i++) {
    long long name, type, normal; int a, b, c;
    in >> name >> a >> b >> c >> type >> normal;
    t->f1.push_back(a); t->f2.push_back(b); t->f3.push_back(c);
  }
  int na = 0;
  in >> d >> d >> d >> d >> d >> d >> na;
  for (int i = 0; i < 5; i++) in >> d;
  for (int i = 0; i < na && in; i++) {
    long long name; int a, b; double eps, sig;
    in >> name >> a >> b >> eps >> sig;
    t->la1.push_back(a); t->la2.push_back(b);
  }
  if (!in || t->np <= 0 || t->ne < 0 || t->nf < 0 || (int)t->x.size() != t->np ||
      (int)t->e1.size() != t->ne || (int)t->f1.size() != t->nf) {
    set_error(std::string("malformed template ") + path);
    delete t;
    return nullptr;
  }
  if (!derive_template(*t)) { delete t; return nullptr; }
  return t;
}

void sf_template_free(sf_template* t) { delete t; }

int sf_template_info(const sf_template* t, int* n_beads, int* n_edges, int* n_faces, int* n_hinges,
                     int* n_charged, int* centre_bead, double* si_sigma, double* si_mass) {
  if (!t) { set_error("null template"); return SF_E_ARG; }
  if (n_beads) *n_beads = t->np;
  if (n_edges) *n_edges = t->ne;
  if (n_faces) *n_faces = t->nf;
  if (n_hinges) *n_hinges = t->nh();
  if (n_charged) *n_charged = t->nq();
  if (centre_bead) *centre_bead = t->centre;
  if (si_sigma) *si_sigma = t->si_sigma;
  if (si_mass) *si_mass = t->si_mass;
  return SF_OK;
}

int sf_template_extent(const sf_template* t, double* diameter, double* r_all, double* r_charged) {
  if (!t) { set_error("null template"); return SF_E_ARG; }
  if (diameter) *diameter = t->diam;
  if (r_all) *r_all = t->r_all;
  if (r_charged) *r_charged = t->r_q;
  return SF_OK;
}

int sf_template_frame(const sf_template* t, int* fp1, int* fp2, double* qbody) {
  if (!t) { set_error("null template"); return SF_E_ARG; }
  if (!t->frame_ok) { set_error("template beads are collinear: no body frame"); return SF_E_ARG; }
  if (fp1) *fp1 = t->fp1;
  if (fp2) *fp2 = t->fp2;
  if (qbody) for (size_t k = 0; k < t->qbody.size(); k++) qbody[k] = t->qbody[k];
  return SF_OK;
}

int sf_template_hinges(const sf_template* t, int* out, int cap) {
  if (!t) { set_error("null template"); return SF_E_ARG; }
  int nh = t->nh();
  for (int i = 0; i < std::min(nh, cap) * 4; i++) out[i] = t->hinges[i];
  return nh;
}

// src/md.cpp:135-173 (box, ecut, lb, ni, kappa with Pi = 3.14159, ecut_el), src/md.cpp:187-196.
int sf_params_derive(const sf_template* t, const sf_cli* c, sf_params* o) {
  if (!t || !c || !o) { set_error("null argument"); return SF_E_ARG; }
  if (c->n_subunits <= 0 || c->C_uM <= 0) { set_error("need n_subunits > 0 and C > 0"); return SF_E_ARG; }
  const double Avagadro = 6.022e23, Pi = 3.14159, qs = 1;
  std::memset(o, 0, sizeof(*o));
  double sig = t->si_sigma;
  o->box = std::pow((double(c->n_subunits) * 1000 / (c->C_uM * std::pow(sig, 3) * Avagadro)), 1.0 / 3.0);
  o->ecut = 2.5 * (sig / 1e-9);
  o->lb = (0.701e-9) / sig;
  double ni = c->c_mM * Avagadro * sig * sig * sig;
  o->kappa = std::sqrt(8 * Pi * ni * o->lb * qs * qs);
  o->ecut_el = (1 / o->kappa) * c->ecut_c;
  o->ks = c->ks; o->kb = c->kb; o->elj_att = c->elj_att;
  o->T = c->K / double(298);
  o->dt = c->dt; o->damp = c->damp; o->Q = c->Q;
  o->q_chain = c->q_chain; o->brownian = c->brownian;
  o->skin = 0.3;
  o->pair_fp32 = 0;
  o->bd_seed = 69521u;
  return SF_OK;
}

// Seeded non-overlapping state. Subunits are inserted one at a time: uniform random centre,
// uniform random rotation (unit quaternion from 4 Gaussians), accepted when every bead is at
// least 1 sigma from every bead already placed (hash grid of 1-sigma cells, minimum image).
int sf_make_state(const sf_template* t, int ns, double box, double T, uint64_t seed, double* pos,
                  double* vel) {
  if (!t || ns <= 0 || !(box > 0) || !pos) { set_error("bad arguments to sf_make_state"); return SF_E_ARG; }
  const int np = t->np;
  const long long N = (long long)ns * np;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-0.5, 0.5);
  std::normal_distribution<double> G(0.0, 1.0);
  double cx = 0, cy = 0, cz = 0;
  for (int l = 0; l < np; l++) { cx += t->x[l]; cy += t->y[l]; cz += t->z[l]; }
  cx /= np; cy /= np; cz /= np;
  const double hbox = 0.5 * box;
  const double dmin2 = 1.0;
  int nc = std::max(1, (int)std::floor(box / 1.0));
  if (nc < 3) nc = 1;
  const double csz = box / nc;
  std::unordered_map<long long, std::vector<int>> grid;
  auto cell_of = [&](double v) {
    int c = (int)std::floor((v + hbox) / csz);
    c %= nc;
    if (c < 0) c += nc;
    return c;
  };
  auto wrap = [&](double v) {
    if (v > hbox) v -= box; else if (v < -hbox) v += box;
    return v;
  };
  auto mi = [&](double v) {
    if (v > hbox) v -= box; else if (v < -hbox) v += box;
    return v;
  };
  std::vector<double> trial(3 * np);
  for (int s = 0; s < ns; s++) {
    int attempts = 0;
    for (;;) {
      if (++attempts > 200000) { set_error("could not place subunits without overlap (box too dense)"); return SF_E_ARG; }
      double c[3] = {U(rng) * box, U(rng) * box, U(rng) * box};
      double q0 = G(rng), q1 = G(rng), q2 = G(rng), q3 = G(rng);
      double qn = std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
      q0 /= qn; q1 /= qn; q2 /= qn; q3 /= qn;
      double R[9] = {q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2),
                     2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1),
                     2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3};
      bool ok = true;
      for (int l = 0; l < np && ok; l++) {
        double bx = t->x[l] - cx, by = t->y[l] - cy, bz = t->z[l] - cz;
        double p[3];
        for (int a = 0; a < 3; a++) p[a] = wrap(c[a] + R[3 * a] * bx + R[3 * a + 1] * by + R[3 * a + 2] * bz);
        trial[3 * l] = p[0]; trial[3 * l + 1] = p[1]; trial[3 * l + 2] = p[2];
        int ci = cell_of(p[0]), cj = cell_of(p[1]), ck = cell_of(p[2]);
        int span = (nc >= 3) ? 1 : 0;
        for (int di = -span; di <= span && ok; di++)
          for (int dj = -span; dj <= span && ok; dj++)
            for (int dk = -span; dk <= span && ok; dk++) {
              long long key = (((long long)((ci + di + nc) % nc) * nc) + (cj + dj + nc) % nc) * nc + (ck + dk + nc) % nc;
              auto it = grid.find(key);
              if (it == grid.end()) continue;
              for (int j : it->second) {
                double dx = mi(p[0] - pos[3 * j]), dy = mi(p[1] - pos[3 * j + 1]), dz = mi(p[2] - pos[3 * j + 2]);
                if (dx * dx + dy * dy + dz * dz < dmin2) { ok = false; break; }
              }
            }
      }
      if (!ok) continue;
      for (int l = 0; l < np; l++) {
        long long g = (long long)s * np + l;
        pos[3 * g] = trial[3 * l]; pos[3 * g + 1] = trial[3 * l + 1]; pos[3 * g + 2] = trial[3 * l + 2];
        long long key = (((long long)cell_of(trial[3 * l]) * nc) + cell_of(trial[3 * l + 1])) * nc + cell_of(trial[3 * l + 2]);
        grid[key].push_back((int)g);
      }
      break;
    }
  }
  if (vel) {
    double mx = 0, my = 0, mz = 0;
    for (long long i = 0; i < N; i++) {
      double sd = std::sqrt(T / t->mass[i % np]);
      vel[3 * i] = G(rng) * sd; vel[3 * i + 1] = G(rng) * sd; vel[3 * i + 2] = G(rng) * sd;
      mx += vel[3 * i]; my += vel[3 * i + 1]; mz += vel[3 * i + 2];
    }
    mx /= N; my /= N; mz /= N;
    double ke = 0;
    for (long long i = 0; i < N; i++) {
      vel[3 * i] -= mx; vel[3 * i + 1] -= my; vel[3 * i + 2] -= mz;
      ke += 0.5 * t->mass[i % np] * (vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] + vel[3 * i + 2] * vel[3 * i + 2]);
    }
    double sc = (ke > 0) ? std::sqrt(1.5 * N * T / ke) : 0.0;
    for (long long i = 0; i < 3 * N; i++) vel[i] *= sc;
  }
  return SF_OK;
}

}  // extern "C"
