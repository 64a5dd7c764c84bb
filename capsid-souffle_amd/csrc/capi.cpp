// C ABI of the SOUFFLE GPU engine: device memory, parameter tables, the step driver with
// list-capacity recovery, energies and state transfer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "sf_far.hpp"

using namespace sf;

namespace {

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                       \
      return SF_E_HIP;                                                                 \
    }                                                                                  \
  } while (0)

constexpr int kChunk = 100;          // steps between capacity checks (default graph length)
constexpr int kMaxGraphSteps = 20000;  // longest graph a handle captures (5 kernel nodes a step)
constexpr int kTimeEvery = 10;       // with timing on, every 10th step's kernels get HIP events
constexpr double kBoundMargin = 1.0; // initial slack of the centre-bead radius (sigma)
constexpr double kNoCut2 = 1e30;     // squared "cutoff" of the uncut Yukawa's bead tests

}  // namespace

// Square of the displacement that triggers a list build. Positions at the last build are kept in
// FP32, so a measured displacement can fall short of the true one by the rounding of three
// coordinates (2^-24 relative each; coordinates stay within a box of the origin); the threshold
// is lowered by twice that bound (a build comes at most that much earlier), so no bead moves
// more than skin/2 between builds and every pair within its cutoff stays in the lists.
static double half_skin_test2(double skin, double box) {
  const double margin = 2.0 * std::sqrt(3.0) * box * std::ldexp(1.0, -24);
  const double h = 0.5 * skin - margin;
  return h > 0 ? h * h : 0.0;
}

struct sf_handle {
  sf_template t;
  int R = 0, Ns = 0, Nb = 0, Nq = 0;
  int dev = 0;                 // HIP device current at sf_create; the handle's memory lives there
  std::vector<int> ns;
  std::vector<sf_params> prm;
  std::vector<RepParams> rp;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  Dev d{};
  double margin = kBoundMargin;
  std::vector<double> far_rf;  // per replica, split mode
  // far-field model (split mode): device copies of its spec and packed weights
  FarDev fdev{};
  bool far_model = false;      // a model is attached and has not failed an audit
  bool far_drop = false;       // far field left out (near field only)
  bool far_fell_back = false;
  double far_tol = 1e-3;
  long long far_audits = 0;
  double far_last = 0, far_max = 0;            // audit error relative to the far field
  double far_last_force = 0, far_max_force = 0;  // the same difference relative to the total force
  double* qfar_exact = nullptr;  // [Nq][3] exact far field at audits
  double* audit_out = nullptr;   // [R][3]
  int mode = MODE_MD;
  int fp32 = 0;
  bool timing = false;
  // steps per graph launch between host checks of the capacity, deformation and finiteness
  // flags (sf_set_graph_steps); launches counted since sf_create
  int graph_steps = kChunk;
  long long graph_launches = 0;
  double t_int = 0, t_rb = 0, t_force = 0, t_red = 0, t_steps = 0, t_far = 0;
  // snapshot for replaying a chunk
  double4* s_pos = nullptr; double* s_v = nullptr; double* s_f = nullptr; double* s_g = nullptr;
  RepState* s_rs = nullptr;
  double* out_buf = nullptr;  // 5 x Nb x 3 for sf_forces
  OligoBuf olig{};            // oligomer analysis buffers, allocated at its first call
  double* tab = nullptr;      // per-replica qq / yk / bd_amp tables

  template <typename T>
  int alloc(T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&q, n * sizeof(T)));
    allocs.push_back(q);
    *p = (T*)q;
    return SF_OK;
  }
  void release(void* p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) { allocs.erase(it); (void)hipFree(p); }
  }
  ~sf_handle() {
    drop_graph();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }

  void fill_bounds() {
    const double rb = t.r_all + margin, rq = t.r_q + margin;
    for (int r = 0; r < R; r++) {
      RepParams& P = rp[r];
      const sf_params& p = prm[r];
      // largest LJ/WCA list radius over the template's bead pairs
      double smax = 0;
      for (int l = 0; l < t.np; l++) smax = std::max(smax, t.sigma[l]);
      bool any_att = false;
      for (int l = 0; l < t.np; l++)
        if (t.att_mask[t.type[l]]) any_att = true;
      double rlj = kWcaFactor * smax;
      if (any_att) rlj = std::max(rlj, p.ecut);
      rlj += p.skin;
      P.rl_lj = rlj;
      // list entries were within their list radius at the build and each bead has moved less
      // than skin/2 since: no entry is farther than the larger list radius plus the skin (the
      // uncut Yukawa's near lists reach far_rf: no bound is used there)
      P.pair_reach = P.split ? 1e300 : std::max(rlj, p.ecut_el + p.skin) + p.skin + 1e-2;
      P.sub_rc_lj2 = (rlj + 2 * rb) * (rlj + 2 * rb);
      const double rql = p.ecut_el + p.skin;
      P.sub_rc_q2 = (rql + 2 * rq) * (rql + 2 * rq);
      // uncut Yukawa: the near lists hold every charged bead of the subunits whose centre
      // beads are within far_rf + skin (centre beads move at most skin/2 between builds)
      if (P.split) P.sub_rc_q2 = (far_rf[r] + p.skin) * (far_rf[r] + p.skin);
      P.r_bound = rb;
      P.r_bound2 = rb * rb;
      P.r_bound_q2 = rq * rq;
      // the LDS build grid spans +-(r_bound + rl_lj) around the centre bead and must stay
      // inside half a box; bead ids must fit the 2^23 float index of the grid records
      P.grid_ok = (rb + rlj + 2e-3 < 0.5 * p.box && (long long)Nb < (1ll << 23)) ? 1 : 0;
    }
  }

  // largest skin replica r can use: minimum image needs max(cutoffs) + skin < box/2
  double max_skin(int r) const {
    const sf_params& p = prm[r];
    return 0.5 * p.box - std::max(std::max(p.ecut, p.ecut_el), far_rf[r]) - 1e-9;
  }

  // Sets replica r's Verlet skin (host tables only; upload_rp and a forced rebuild follow).
  void put_skin(int r, double skin) {
    prm[r].skin = skin;
    RepParams& P = rp[r];
    P.skin = skin;
    P.half_skin2 = half_skin_test2(skin, P.box);
    P.rc_q_list2 = P.split ? kNoCut2 : (prm[r].ecut_el + skin) * (prm[r].ecut_el + skin);
  }

  // Marks every replica for a list build at the next step.
  int request_rebuild() {
    std::vector<RepState> h;
    if (int rc = read_states(h)) return rc;
    for (auto& s : h) s.rebuild = 1;
    HIPCHK(hipMemcpyAsync(d.rs, h.data(), R * sizeof(RepState), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return SF_OK;
  }

  // far field of the uncut Yukawa (split mode only): the model when one is attached (energies
  // always use the exact far field)
  int launch_far(int energy) {
    if (d.nbig && energy)  // big-bead energies; their forces are evaluated inside k_force
      if (int rc = launch_big(d, st)) return rc;
    if (!d.split) return SF_OK;
    if (far_drop && !energy) return launch_fill(d.qfar, 3 * (size_t)d.Nq, 0.0, st);
    if (far_model && !energy) {
      if (int rc = launch_frames(d, fdev, st)) return rc;
      return launch_far_mlp(d, fdev, st);
    }
    return launch_far_exact(d, energy, st);
  }

  // relative error of the model's far field (in d.qfar after the last force evaluation)
  // against the exact one at the current positions: per replica
  // sqrt(sum |F_model - F_exact|^2 / sum |F_exact|^2) over its charged beads, i.e. relative to
  // the far field the model replaces; the largest over the replicas
  int audit(double* err) {
    Dev dx = d;
    dx.qfar = qfar_exact;
    if (int rc = launch_far_exact(dx, 0, st)) return rc;
    if (int rc = launch_far_audit(d, qfar_exact, audit_out, st)) return rc;
    std::vector<double> h3(3 * (size_t)R);
    HIPCHK(hipMemcpyAsync(h3.data(), audit_out, h3.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double worst = 0.0, worst_f = 0.0;
    auto ratio = [](double a, double b) { return b > 0 ? std::sqrt(a / b) : (a > 0 ? HUGE_VAL : 0.0); };
    for (int r = 0; r < R; r++) {
      const double e = ratio(h3[3 * r], h3[3 * r + 1]), ef = ratio(h3[3 * r], h3[3 * r + 2]);
      if (!(e <= worst)) worst = e;  // a NaN error counts as the worst
      if (!(ef <= worst_f)) worst_f = ef;
    }
    *err = worst;
    far_audits++;
    far_last = worst;
    far_max = std::max(far_max, worst);
    far_last_force = worst_f;
    far_max_force = std::max(far_max_force, worst_f);
    return SF_OK;
  }

  // Cell grids of the list build (cut Yukawa): per replica g cells per axis with edge at least
  // the list radius (+2e-3 for the FP32 tests), g = 1 when fewer than 3 fit; global cell
  // arrays sized for all replicas.
  int setup_cells() {
    if (!d.cells) return SF_OK;
    long long nlj = 0, nq = 0;
    for (int r = 0; r < R; r++) {
      RepParams& P = rp[r];
      // at most about two beads per cell on average: finer grids only cost memory
      auto grid = [&](double radius, long long n) {
        int g = (int)std::floor(P.box / (radius + 2e-3));
        const int gmax = (int)std::floor(std::cbrt(2.0 * (double)std::max(n, 1ll)));
        g = std::min(g, gmax);
        return g < 3 ? 1 : g;
      };
      P.g_lj = grid(P.rl_lj, (long long)P.nb);
      P.g_q = grid(prm[r].ecut_el + prm[r].skin, (long long)P.nqr);
      P.inv_a_lj = (float)(P.g_lj / P.box);
      P.inv_a_q = (float)(P.g_q / P.box);
      P.c0_lj = (int)nlj;
      P.c0_q = (int)nq;
      nlj += (long long)P.g_lj * P.g_lj * P.g_lj;
      nq += (long long)P.g_q * P.g_q * P.g_q;
    }
    if (nlj > (1ll << 30) || nq > (1ll << 30)) { set_error("too many list-build cells"); return SF_E_ARG; }
    if (nlj != d.ncell_lj) {
      release(d.cnt_lj);
      if (alloc(&d.cnt_lj, (size_t)nlj)) return SF_E_HIP;
      d.ncell_lj = (int)nlj;
      const int want = std::max(16, (int)(4.0 * Nb / nlj) + 4);
      if (int rc = grow_cells(std::max(want, d.ccap_lj), 0, true)) return rc;
      HIPCHK(hipMemsetAsync(d.cnt_lj, 0, (size_t)nlj * sizeof(int), st));
    }
    if (nq != d.ncell_q) {
      release(d.cnt_q);
      if (alloc(&d.cnt_q, (size_t)nq)) return SF_E_HIP;
      d.ncell_q = (int)nq;
      const int want = std::max(16, (int)(4.0 * std::max(Nq, 1) / nq) + 4);
      if (int rc = grow_cells(0, std::max(want, d.ccap_q), true)) return rc;
      HIPCHK(hipMemsetAsync(d.cnt_q, 0, (size_t)nq * sizeof(int), st));
    }
    return SF_OK;
  }

  // (Re)allocates the cell records for capacities of at least lj / q slots per cell.
  int grow_cells(int lj, int q, bool exact = false) {
    if (lj > d.ccap_lj || (exact && lj > 0)) {
      const int cap = exact ? lj : std::max(lj + lj / 4 + 4, 2 * d.ccap_lj);
      release(d.cell_lj);
      if (alloc(&d.cell_lj, (size_t)d.ncell_lj * cap)) return SF_E_HIP;
      d.ccap_lj = cap;
    }
    if (q > d.ccap_q || (exact && q > 0)) {
      const int cap = exact ? q : std::max(q + q / 4 + 4, 2 * d.ccap_q);
      release(d.cell_q);
      if (alloc(&d.cell_q, (size_t)std::max(d.ncell_q, 1) * cap)) return SF_E_HIP;
      d.ccap_q = cap;
    }
    return SF_OK;
  }

  int upload_rp() {
    HIPCHK(hipMemcpyAsync((void*)d.rp, rp.data(), R * sizeof(RepParams), hipMemcpyHostToDevice, st));
    return SF_OK;
  }

  int grow_lists(int need_lj, int need_q) {
    if (need_lj > d.lj_cap) {
      int cap = std::max(need_lj + need_lj / 4 + 8, d.lj_cap * 2);
      release(d.lj_list);
      int* p;
      if (alloc(&p, (size_t)cap * Nb)) return SF_E_HIP;
      d.lj_list = p; d.lj_cap = cap;
    }
    if (need_q > d.q_cap) {
      int cap = std::max(need_q + need_q / 4 + 8, d.q_cap * 2);
      release(d.q_list);
      int* p;
      if (alloc(&p, (size_t)cap * std::max(Nq, 1))) return SF_E_HIP;
      d.q_list = p; d.q_cap = cap;
    }
    return SF_OK;
  }

  int read_states(std::vector<RepState>& h) {
    h.resize(R);
    HIPCHK(hipMemcpyAsync(h.data(), d.rs, R * sizeof(RepState), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return SF_OK;
  }

  // After a launch sequence: inspect non-finite, overflow and deformation flags. Returns 2 if a
  // replica became non-finite (bad_step, bad_rep: the first such step and its replica; nothing
  // is regrown, since a non-finite state can ask for any capacity), 1 if the sequence must be
  // replayed (capacities or bounds were raised), 0 if fine, <0 on error.
  long long bad_step = 0;
  int bad_rep = -1;
  int check_and_fix() {
    std::vector<RepState> h;
    if (int rc = read_states(h)) return rc;
    bad_rep = -1;
    for (int r = 0; r < R; r++)
      if (h[r].nonfinite && (bad_rep < 0 || h[r].bad_step < bad_step)) { bad_rep = r; bad_step = h[r].bad_step; }
    if (bad_rep >= 0) return 2;
    int need_lj = 0, need_q = 0, need_c = 0, deformed = 0, need_clj = 0, need_cq = 0;
    for (auto& s : h) {
      need_lj = std::max(need_lj, s.need_lj);
      need_q = std::max(need_q, s.need_q);
      need_c = std::max(need_c, s.need_cand);
      need_clj = std::max(need_clj, s.need_cell_lj);
      need_cq = std::max(need_cq, s.need_cell_q);
      deformed |= s.deformed;
    }
    if (!need_lj && !need_q && !need_c && !deformed && !need_clj && !need_cq) return 0;
    if (need_clj || need_cq)
      if (int rc = grow_cells(need_clj, need_cq)) return rc;
    if (need_c) {
      set_error("more than 1024 candidate subunits near one subunit at a list build");
      return SF_E_OVERFLOW;
    }
    if (deformed) {
      margin += 1.0;
      if (margin > 20.0) { set_error("subunit deformed beyond 20 sigma of its centre bead"); return SF_E_OVERFLOW; }
      fill_bounds();
      if (int rc = upload_rp()) return rc;
    }
    if (int rc = grow_lists(need_lj, need_q)) return rc;
    return 1;
  }

  int clear_flags_force_rebuild() {
    std::vector<RepState> h;
    if (int rc = read_states(h)) return rc;
    for (auto& s : h) { s.need_lj = s.need_q = s.need_cand = s.deformed = s.need_cell_lj = s.need_cell_q = 0; s.rebuild = 1; }
    HIPCHK(hipMemcpyAsync(d.rs, h.data(), R * sizeof(RepState), hipMemcpyHostToDevice, st));
    return SF_OK;
  }

  int snapshot(bool save) {
    const size_t nb = Nb;
    hipMemcpyKind k = hipMemcpyDeviceToDevice;
    if (save) {
      HIPCHK(hipMemcpyAsync(s_pos, d.pos, nb * sizeof(double4), k, st));
      HIPCHK(hipMemcpyAsync(s_v, d.vx, 3 * nb * sizeof(double), k, st));
      HIPCHK(hipMemcpyAsync(s_f, d.fx, 3 * nb * sizeof(double), k, st));
      HIPCHK(hipMemcpyAsync(s_g, d.gx, 3 * nb * sizeof(double), k, st));
      HIPCHK(hipMemcpyAsync(s_rs, d.rs, R * sizeof(RepState), k, st));
    } else {
      HIPCHK(hipMemcpyAsync(d.pos, s_pos, nb * sizeof(double4), k, st));
      HIPCHK(hipMemcpyAsync(d.vx, s_v, 3 * nb * sizeof(double), k, st));
      HIPCHK(hipMemcpyAsync(d.fx, s_f, 3 * nb * sizeof(double), k, st));
      HIPCHK(hipMemcpyAsync(d.gx, s_g, 3 * nb * sizeof(double), k, st));
      HIPCHK(hipMemcpyAsync(d.rs, s_rs, R * sizeof(RepState), k, st));
      if (int rc = launch_centres(d, st)) return rc;
    }
    return SF_OK;
  }

  // start: forces with a list build, kinetic energy, thermostat and noise set-up
  int start_sequence() {
    for (int attempt = 0; attempt < 16; attempt++) {
      if (int rc = launch_centres(d, st)) return rc;
      if (int rc = launch_rebuild(d, 1, st)) return rc;
      if (int rc = launch_far(0)) return rc;
      if (int rc = launch_force(d, MODE_FORCES, 1, fp32, st)) return rc;
      if (int rc = launch_finalize(d, 1, st)) return rc;
      int rc = check_and_fix();
      if (rc < 0) return rc;
      if (rc == 0) return SF_OK;
      if (rc == 2) {
        set_error("non-finite initial state (replica " + std::to_string(bad_rep) + ")");
        return SF_E_NONFINITE;
      }
      std::vector<RepState> h;
      if ((rc = read_states(h))) return rc;
      for (auto& s : h) { s.need_lj = s.need_q = s.need_cand = s.deformed = s.need_cell_lj = s.need_cell_q = 0; }
      HIPCHK(hipMemcpyAsync(d.rs, h.data(), R * sizeof(RepState), hipMemcpyHostToDevice, st));
    }
    set_error("list capacity did not converge");
    return SF_E_OVERFLOW;
  }

  // n steps of the loop body: integrate, (conditional) list build, forces, reductions. With
  // timing on, every kTimeEvery-th step's kernels are bracketed by events (kEv per sample).
  static constexpr int kEv = 6;
  int run_chunk(int n, bool tm_on) {
    for (int k = 0; k < n; k++) {
      const bool tm = tm_on && (k % kTimeEvery == 0);
      hipEvent_t* e = tm ? &events[kEv * (k / kTimeEvery)] : nullptr;
      if (tm) HIPCHK(hipEventRecord(e[0], st));
      if (int rc = launch_integrate(d, mode, st)) return rc;
      if (tm) HIPCHK(hipEventRecord(e[1], st));
      if (int rc = launch_rebuild(d, 0, st)) return rc;
      if (tm) HIPCHK(hipEventRecord(e[2], st));
      if (int rc = launch_far(0)) return rc;
      if (tm) HIPCHK(hipEventRecord(e[3], st));
      if (int rc = launch_force(d, mode, 0, fp32, st)) return rc;
      if (tm) HIPCHK(hipEventRecord(e[4], st));
      if (int rc = launch_finalize(d, 0, st)) return rc;
      if (tm) HIPCHK(hipEventRecord(e[5], st));
    }
    return SF_OK;
  }

  // Each chunk length is captured once into a HIP graph and replayed (a few lengths are kept:
  // the full chunk and the remainders callers use); every capture is redone when the timing
  // mode, the tile, the far-field form or any list buffer (reallocated on growth) changes.
  static constexpr int kGraphCache = 4;
  std::vector<std::pair<int, hipGraphExec_t>> graphs;  // (chunk length, executable), oldest first
  bool g_valid = false;
  bool g_tm = false;
  const void* g_key[2] = {nullptr, nullptr};
  int g_cap[2] = {0, 0};
  int g_tile = 0;
  bool g_far = false, g_drop = false;
  int g_nbig = 0;
  // encapsulation big beads: buffers (allocated on first use) and whether the set changed
  // since the last start
  int nbig_cap = 0;
  double4* big_pos = nullptr;
  double4* big_tab = nullptr;
  double4* big_f = nullptr;
  bool big_dirty = false;
  const void* g_cell[4] = {nullptr, nullptr, nullptr, nullptr};
  int g_ccap[2] = {0, 0};
  std::vector<hipEvent_t> events;

  // The executable graph's last launch may still be running on the stream: it is drained before
  // the graph is destroyed (HIP does not promise to defer the destruction of a running graph).
  void drop_graph() {
    if (!graphs.empty()) (void)hipStreamSynchronize(st);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
    g_valid = false;
  }
  int ensure_events() {
    const size_t need = kEv * ((kChunk + kTimeEvery - 1) / kTimeEvery);
    while (events.size() < need) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      events.push_back(e);
    }
    return SF_OK;
  }
  int launch_chunk(int n) {
    if (timing) {
      // HIP does not take timing events into a captured graph: timed chunks launch directly
      if (int rc = ensure_events()) return rc;
      return run_chunk(n, true);
    }
    // the captured kernel arguments hold the list buffers and their capacities; a regrown
    // buffer may come back at the same address, so the capacities are part of the key
    if (!g_valid || g_tm != timing || g_key[0] != d.lj_list || g_key[1] != d.q_list ||
        g_cap[0] != d.lj_cap || g_cap[1] != d.q_cap || g_tile != d.tile || g_far != far_model || g_nbig != d.nbig ||
        g_drop != far_drop || g_cell[0] != d.cell_lj || g_cell[1] != d.cnt_lj || g_cell[2] != d.cell_q ||
        g_cell[3] != d.cnt_q || g_ccap[0] != d.ccap_lj || g_ccap[1] != d.ccap_q) {
      drop_graph();
      g_tm = timing; g_key[0] = d.lj_list; g_key[1] = d.q_list;
      g_cap[0] = d.lj_cap; g_cap[1] = d.q_cap; g_tile = d.tile; g_far = far_model; g_drop = far_drop; g_nbig = d.nbig;
      g_cell[0] = d.cell_lj; g_cell[1] = d.cnt_lj; g_cell[2] = d.cell_q; g_cell[3] = d.cnt_q;
      g_ccap[0] = d.ccap_lj; g_ccap[1] = d.ccap_q;
      g_valid = true;
    }
    hipGraphExec_t ex = nullptr;
    for (auto& g : graphs)
      if (g.first == n) ex = g.second;
    if (!ex) {
      if ((int)graphs.size() >= kGraphCache) {  // the oldest length goes (after its last launch)
        HIPCHK(hipStreamSynchronize(st));
        (void)hipGraphExecDestroy(graphs.front().second);
        graphs.erase(graphs.begin());
      }
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = run_chunk(n, timing);
      hipError_t ec = hipStreamEndCapture(st, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (ec != hipSuccess) { set_error(std::string("graph capture failed: ") + hipGetErrorString(ec)); return SF_E_HIP; }
      hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) { set_error(std::string("graph instantiate failed: ") + hipGetErrorString(ei)); return SF_E_HIP; }
      graphs.emplace_back(n, ex);
    }
    HIPCHK(hipGraphLaunch(ex, st));
    graph_launches++;
    return SF_OK;
  }
};

extern "C" {

sf_handle* sf_create(const sf_template* t, int R, const int* n_subunits, const sf_params* params) {
  if (!t || R <= 0 || !n_subunits || !params) { set_error("bad arguments to sf_create"); return nullptr; }
  const int np = t->np, nq = t->nq();
  if (np > kBlock) { set_error("template has more than 128 beads"); return nullptr; }
  if ((size_t)(2 * np) * sizeof(double4) + (size_t)t->nh() * 12 * sizeof(double) > 60 * 1024) {
    set_error("template too large for the subunit workgroup's LDS"); return nullptr;
  }
  int brown = params[0].brownian;
  int fp32 = params[0].pair_fp32;
  const int split = params[0].ecut_el == 0 ? 1 : 0;
  long long Ns = 0;
  for (int r = 0; r < R; r++) {
    const sf_params& p = params[r];
    if (n_subunits[r] <= 0) { set_error("replica with no subunits"); return nullptr; }
    if (p.brownian != brown || p.pair_fp32 != fp32) { set_error("all replicas must share the integrator and pair precision"); return nullptr; }
    if (!(p.box > 0) || !(p.dt > 0) || !(p.skin > 0) || !(p.kappa > 0)) { set_error("box, dt, skin and kappa must be positive"); return nullptr; }
    if (!(p.ecut_el >= 0)) { set_error("ecut_el must be >= 0"); return nullptr; }
    if ((p.ecut_el == 0) != (split == 1)) { set_error("all replicas must share the Yukawa form (cut or uncut)"); return nullptr; }
    if (split) {
      const double rf = p.far_rf > 0 ? p.far_rf : 2.5 * t->diam;
      if (!(p.far_rf >= 0) || rf + p.skin >= 0.5 * p.box) {
        set_error("uncut Yukawa: far_rf + skin must be below half the box"); return nullptr;
      }
      if (nq > 8) { set_error("uncut Yukawa: the far field supports at most 8 charged beads per template"); return nullptr; }
    }
    if (!brown && (p.q_chain < 1 || p.q_chain > kMaxChain)) { set_error("Nose-Hoover chain length must be 1..8"); return nullptr; }
    if (brown && !(p.damp > 0)) { set_error("BD damping must be positive"); return nullptr; }
    double reach = std::max(p.ecut, p.ecut_el) + p.skin;
    if (reach > 0.5 * p.box) {
      set_error("box too small for the cutoffs (minimum image needs cutoff + skin < box/2)"); return nullptr;
    }
    Ns += n_subunits[r];
  }
  if (Ns * np > 0x7fffffffLL / 2) { set_error("too many beads"); return nullptr; }
  sf_handle* h = new sf_handle();
  h->t = *t;
  h->R = R;
  h->Ns = (int)Ns;
  h->Nb = (int)Ns * np;
  h->Nq = (int)Ns * nq;
  h->ns.assign(n_subunits, n_subunits + R);
  h->prm.assign(params, params + R);
  h->mode = brown ? MODE_BD : MODE_MD;
  h->fp32 = fp32;
  h->far_rf.assign(R, 0.0);
  for (int r = 0; r < R; r++) h->far_rf[r] = split ? (params[r].far_rf > 0 ? params[r].far_rf : 2.5 * t->diam) : 0.0;
  auto fail = [&](const char* what) -> sf_handle* {
    if (what) set_error(what);
    delete h;
    return nullptr;
  };
  if (hipGetDevice(&h->dev) != hipSuccess) return fail("hipGetDevice failed (no GPU?)");
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) return fail("hipStreamCreate failed (no GPU?)");
  Dev& d = h->d;
  d.Nb = h->Nb; d.Ns = h->Ns; d.Nq = h->Nq; d.R = R;
  d.tile = 384;  // the autotuner may choose another
  d.split = split;
  // template tables
  TopoDev& td = d.topo;
  td.np = np; td.nh = t->nh(); td.nq = nq; td.centre = t->centre;
  td.fp1 = t->fp1; td.fp2 = t->fp2;
  auto up = [&](auto** dst, const auto& v) -> int {
    typedef typename std::remove_reference<decltype(v[0])>::type T0;
    typedef typename std::remove_const<T0>::type T;
    T* p;
    if (h->alloc(&p, v.size())) return SF_E_HIP;
    if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("hipMemcpy failed"); return SF_E_HIP;
    }
    *dst = p;
    return SF_OK;
  };
  // padded per-bead rows of incident edges and hinge incidences
  std::vector<int> e_cnt(np), h_cnt(np);
  std::vector<unsigned char> e_oth((size_t)np * kMaxDeg, 0);
  std::vector<unsigned short> h_sl((size_t)np * kMaxHinc, 0);
  std::vector<double> e_l0((size_t)np * kMaxDeg, 0.0);
  if (4 * t->nh() > 65535) { set_error("template has too many bending hinges"); delete h; return nullptr; }
  for (int l = 0; l < np; l++) {
    e_cnt[l] = t->e_off[l + 1] - t->e_off[l];
    h_cnt[l] = t->h_off[l + 1] - t->h_off[l];
    if (e_cnt[l] > kMaxDeg || h_cnt[l] > kMaxHinc) {
      set_error("template bead with more than 8 edges or 16 hinge incidences");
      delete h;
      return nullptr;
    }
    for (int k = 0; k < e_cnt[l]; k++) {
      e_oth[(size_t)l * kMaxDeg + k] = t->e_other[t->e_off[l] + k];
      e_l0[(size_t)l * kMaxDeg + k] = t->e_len0[t->e_off[l] + k];
    }
    for (int k = 0; k < h_cnt[l]; k++) h_sl[(size_t)l * kMaxHinc + k] = t->h_slot[t->h_off[l] + k];
  }
  std::vector<int> hinges = t->hinges;
  if (hinges.empty()) hinges.assign(4, 0);
  if (up((double**)&td.mass, t->mass) || up((double**)&td.sigma, t->sigma) || up((int**)&td.type, t->type) ||
      up((int**)&td.qslot, t->qslot) || up((int**)&td.qlocal, t->qlocal) || up((int**)&td.e_cnt, e_cnt) ||
      up((unsigned char**)&td.e_other, e_oth) || up((double**)&td.e_len0, e_l0) || up((int**)&td.h_cnt, h_cnt) ||
      up((unsigned short**)&td.h_slot, h_sl) || up((int**)&td.hinge, hinges) ||
      up((unsigned long long**)&td.att_mask, t->att_mask))
    return fail(nullptr);
  // per-replica tables: qq [nq*nq], yk [nq*nq], bd_amp [np]
  const size_t tstride = 2 * (size_t)nq * nq + np;
  std::vector<double> tabs(tstride * R);
  h->rp.assign(R, RepParams{});
  std::vector<int> sub_rep(Ns);
  int s0 = 0;
  for (int r = 0; r < R; r++) {
    const sf_params& p = params[r];
    double* qq = &tabs[tstride * r];
    double* yk = qq + (size_t)nq * nq;
    double* amp = yk + (size_t)nq * nq;
    for (int a = 0; a < nq; a++)
      for (int b = 0; b < nq; b++) {
        int la = t->qlocal[a], lb = t->qlocal[b];
        double sig1 = t->sigma[la], sig2 = t->sigma[lb];
        double shc = (sig1 + sig2) / 2;
        qq[a * nq + b] = t->q[la] * t->q[lb] * p.lb;
        yk[a * nq + b] = std::exp(p.kappa * shc) / ((1 + p.kappa * sig1 / 2) * (1 + p.kappa * sig2 / 2));
      }
    for (int l = 0; l < np; l++) amp[l] = brown ? std::sqrt((t->mass[l] * 24.0) / (p.damp * p.dt)) : 0.0;
    RepParams& P = h->rp[r];
    P.box = p.box; P.hbox = p.box * 0.5; P.ibox = 1.0 / p.box;
    P.ecut = p.ecut; P.ecut2 = p.ecut * p.ecut;
    P.ecut_el = p.ecut_el; P.ecut_el2 = p.ecut_el * p.ecut_el;
    P.kappa = p.kappa; P.lb = p.lb; P.elj_att = p.elj_att;
    P.ks = p.ks; P.kb = p.kb; P.dt = p.dt; P.damp = p.damp; P.T = p.T;
    P.skin = p.skin; P.half_skin2 = half_skin_test2(p.skin, P.box);
    P.rc_q_list2 = (p.ecut_el + p.skin) * (p.ecut_el + p.skin);
    P.split = split;
    P.far_rf2 = h->far_rf[r] * h->far_rf[r];
    if (split) { P.ecut_el2 = kNoCut2; P.rc_q_list2 = kNoCut2; }
    P.s0 = s0; P.ns = n_subunits[r]; P.b0 = s0 * np; P.nb = n_subunits[r] * np;
    P.q0 = s0 * nq; P.nqr = n_subunits[r] * nq;
    P.inv_nb = 1.0 / (double)P.nb;
    P.q_chain = brown ? 0 : p.q_chain;
    P.brownian = brown; P.pair_fp32 = fp32;
    P.seed = p.bd_seed % kLcgMod;
    if (P.seed == 0) P.seed = 1;  // minstd_rand0 maps seed 0 to 1
    P.mult6n = lcg_pow(6ull * (unsigned long long)P.nb);
    for (int j = 0; j < kMaxChain; j++) { P.Qb[j] = 0; P.dof[j] = 0; }
    if (!brown) {
      const double dof3n = (double)(3ul * (unsigned long)P.nb);
      if (p.q_chain == 1) { P.Qb[0] = 0; P.dof[0] = dof3n; }
      else {
        P.Qb[0] = p.Q; P.dof[0] = dof3n;
        for (int j = 1; j < p.q_chain - 1; j++) { P.Qb[j] = 1; P.dof[j] = 1; }
        P.Qb[p.q_chain - 1] = 0; P.dof[p.q_chain - 1] = dof3n;
      }
    }
    for (int j = 0; j < kMaxChain; j++) P.inv_Qb[j] = P.Qb[j] != 0 ? 1.0 / P.Qb[j] : 0.0;
    for (int s = 0; s < n_subunits[r]; s++) sub_rep[s0 + s] = r;
    s0 += n_subunits[r];
  }
  if (up(&h->tab, tabs)) return fail(nullptr);
  for (int r = 0; r < R; r++) {
    h->rp[r].qq = h->tab + tstride * r;
    h->rp[r].yk = h->tab + tstride * r + (size_t)nq * nq;
    h->rp[r].bd_amp = h->tab + tstride * r + 2 * (size_t)nq * nq;
  }
  h->fill_bounds();
  RepParams* rpd;
  RepState* rsd;
  int* srd;
  if (h->alloc(&rpd, R) || h->alloc(&rsd, R) || h->alloc(&srd, Ns)) return fail(nullptr);
  d.rp = rpd; d.rs = rsd; d.sub_rep = srd;
  if (hipMemcpy(srd, sub_rep.data(), Ns * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return fail("hipMemcpy failed");
  if (hipMemcpy(rpd, h->rp.data(), R * sizeof(RepParams), hipMemcpyHostToDevice) != hipSuccess) return fail("hipMemcpy failed");
  const size_t nb = h->Nb;
  double* vbuf; double* fbuf; double* gbuf;
  if (h->alloc(&d.pos, nb) || h->alloc(&vbuf, 3 * nb) || h->alloc(&fbuf, 3 * nb) || h->alloc(&gbuf, 3 * nb) ||
      h->alloc(&d.xb, nb) || h->alloc(&d.centres, Ns) || h->alloc(&d.lj_cnt, nb) ||
      h->alloc(&d.q_cnt, std::max(h->Nq, 1)) || h->alloc(&d.part, (size_t)Ns * kPartStride) ||
      h->alloc(&d.epart, (size_t)Ns * 5) || h->alloc(&d.ebead, nb * 5) ||
      h->alloc(&h->s_pos, nb) || h->alloc(&h->s_v, 3 * nb) || h->alloc(&h->s_f, 3 * nb) ||
      h->alloc(&h->s_g, 3 * nb) || h->alloc(&h->s_rs, R) || h->alloc(&d.qpos, std::max(h->Nq, 1)) ||
      h->alloc(&d.qfar, 3 * (size_t)std::max(h->Nq, 1)) || h->alloc(&d.efar, std::max(h->Nq, 1)))
    return fail(nullptr);
  d.vx = vbuf; d.vy = vbuf + nb; d.vz = vbuf + 2 * nb;
  d.fx = fbuf; d.fy = fbuf + nb; d.fz = fbuf + 2 * nb;
  d.gx = gbuf; d.gy = gbuf + nb; d.gz = gbuf + 2 * nb;
  d.bd_skip = nullptr;
  if (brown) {  // skip-ahead multipliers of the BD engine, one per replica-local bead
    int max_nb = 0;
    for (int r = 0; r < R; r++) max_nb = std::max(max_nb, h->rp[r].nb);
    std::vector<uint32_t> skip(max_nb);
    const uint64_t m6 = lcg_pow(6);
    uint64_t m = 1;
    for (int b = 0; b < max_nb; b++) {
      skip[b] = (uint32_t)m;
      m = (m * m6) % kLcgMod;
    }
    uint32_t* sk;
    if (h->alloc(&sk, max_nb) ||
        hipMemcpy(sk, skip.data(), max_nb * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
      return fail(nullptr);
    d.bd_skip = sk;
  }
  d.lj_cap = 0; d.q_cap = 0;
  if (h->grow_lists(24, nq > 0 ? 48 : 1)) return fail(nullptr);
  d.cells = (!split && (long long)h->Nb < (1ll << 23)) ? 1 : 0;
  if (h->setup_cells() || h->upload_rp()) return fail(nullptr);
  if (hipMemset(d.qfar, 0, 3 * (size_t)std::max(h->Nq, 1) * sizeof(double)) != hipSuccess ||
      hipMemset(d.efar, 0, (size_t)std::max(h->Nq, 1) * sizeof(double)) != hipSuccess ||
      hipMemset(fbuf, 0, 3 * nb * sizeof(double)) != hipSuccess || hipMemset(gbuf, 0, 3 * nb * sizeof(double)) != hipSuccess ||
      hipMemset(rsd, 0, R * sizeof(RepState)) != hipSuccess)
    return fail("hipMemset failed");
  return h;
}

void sf_destroy(sf_handle* h) { delete h; }

}  // extern "C"

namespace {

// Positions [N][3] as the device's double4 records (x, y, z, tag).
std::vector<double4> tagged_positions(const sf_handle* h, const double* pos) {
  const int np = h->t.np;
  const size_t nb = h->Nb;
  std::vector<double4> hp(nb);
  for (size_t i = 0; i < nb; i++) {
    int s = (int)(i / np), l = (int)(i % np);
    hp[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], make_tag(s, l));
  }
  return hp;
}

// Records per replica whether every given coordinate lies inside [-box/2, box/2] (the steps keep
// it so: one periodic shift per step), which lets the force kernel skip the image shift of list
// neighbours of beads away from the box faces; uploads the replica constants.
int note_wrapped(sf_handle* h, const double* pos) {
  const int np = h->t.np;
  for (int r = 0; r < h->R; r++) {
    RepParams& P = h->rp[r];
    bool in = true;
    for (size_t i = (size_t)P.s0 * np; in && i < (size_t)(P.s0 + P.ns) * np; i++)
      for (int c = 0; c < 3; c++)
        if (!(std::fabs(pos[3 * i + c]) <= P.hbox)) { in = false; break; }
    P.wrapped = in ? 1 : 0;
  }
  return h->upload_rp();
}

// Uploads a state and starts every replica: from scratch (res == nullptr: thermostat zero, BD
// engine seeded, step 0) or from a resume record per replica.
int set_state_impl(sf_handle* h, const double* pos, const double* vel, const sf_resume* res, const double* tforce) {
  const size_t nb = h->Nb;
  std::vector<double4> hp = tagged_positions(h, pos);
  std::vector<double> hv(3 * nb, 0.0);
  if (vel)
    for (size_t i = 0; i < nb; i++) {
      hv[i] = vel[3 * i]; hv[nb + i] = vel[3 * i + 1]; hv[2 * nb + i] = vel[3 * i + 2];
    }
  std::vector<RepState> rs(h->R);
  std::memset(rs.data(), 0, h->R * sizeof(RepState));
  for (int r = 0; r < h->R; r++) {
    rs[r].rng = h->rp[r].seed;
    rs[r].expfac = 1.0;
    if (!res) continue;
    const sf_resume& x = res[r];
    const RepParams& P = h->rp[r];
    if (x.step < 0 || (!P.brownian && x.q_chain != P.q_chain)) {
      set_error("resume record " + std::to_string(r) + " does not fit the replica (step or chain length)");
      return SF_E_ARG;
    }
    if (P.brownian && (x.rng == 0 || x.rng >= kLcgMod)) {
      set_error("resume record " + std::to_string(r) + ": BD engine state must be in 1 .. 2^31-2");
      return SF_E_ARG;
    }
    if (!std::isfinite(x.ke)) { set_error("resume record with a non-finite kinetic energy"); return SF_E_ARG; }
    rs[r].step = x.step;
    if (P.brownian) rs[r].rng = x.rng;
    for (int j = 0; j < P.q_chain; j++) { rs[r].xi[j] = x.xi[j]; rs[r].eta[j] = x.eta[j]; }
    rs[r].ke = x.ke;
    rs[r].resume = 1;
  }
  HIPCHK(hipMemcpyAsync(h->d.pos, hp.data(), nb * sizeof(double4), hipMemcpyHostToDevice, h->st));
  if (int rc = note_wrapped(h, pos)) return rc;
  HIPCHK(hipMemcpyAsync(h->d.vx, hv.data(), 3 * nb * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemsetAsync(h->d.fx, 0, 3 * nb * sizeof(double), h->st));
  HIPCHK(hipMemsetAsync(h->d.gx, 0, 3 * nb * sizeof(double), h->st));
  HIPCHK(hipMemcpyAsync(h->d.rs, rs.data(), h->R * sizeof(RepState), hipMemcpyHostToDevice, h->st));
  int rc = h->start_sequence();
  if (rc) return rc;
  if (tforce) {  // the integrator force of the interrupted run (BD: with its noise and drag)
    std::vector<double> hf(3 * nb);
    for (size_t i = 0; i < nb; i++) {
      hf[i] = tforce[3 * i]; hf[nb + i] = tforce[3 * i + 1]; hf[2 * nb + i] = tforce[3 * i + 2];
    }
    HIPCHK(hipMemcpyAsync(h->d.fx, hf.data(), 3 * nb * sizeof(double), hipMemcpyHostToDevice, h->st));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  h->big_dirty = false;
  return SF_OK;
}

}  // namespace

extern "C" {

int sf_set_state(sf_handle* h, const double* pos, const double* vel) {
  if (!h || !pos) { set_error("null argument"); return SF_E_ARG; }
  return set_state_impl(h, pos, vel, nullptr, nullptr);
}

int sf_set_state_resume(sf_handle* h, const double* pos, const double* vel, const sf_resume* res,
                        const double* tforce) {
  if (!h || !pos || !vel || !res) { set_error("null argument"); return SF_E_ARG; }
  if (h->mode == MODE_BD && !tforce) { set_error("a BD resume needs the integrator force (tforce)"); return SF_E_ARG; }
  return set_state_impl(h, pos, vel, res, tforce);
}

int sf_get_resume(sf_handle* h, sf_resume* out, double* tforce) {
  if (!h || !out) { set_error("null argument"); return SF_E_ARG; }
  std::vector<RepState> s;
  if (int rc = h->read_states(s)) return rc;
  for (int r = 0; r < h->R; r++) {
    sf_resume& x = out[r];
    std::memset(&x, 0, sizeof(x));
    x.step = s[r].step;
    x.q_chain = h->rp[r].q_chain;
    x.rng = s[r].rng;
    x.ke = s[r].ke;
    for (int j = 0; j < h->rp[r].q_chain; j++) { x.xi[j] = s[r].xi[j]; x.eta[j] = s[r].eta[j]; }
  }
  if (tforce) {
    const size_t nb = h->Nb;
    std::vector<double> hf(3 * nb);
    HIPCHK(hipMemcpyAsync(hf.data(), h->d.fx, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t i = 0; i < nb; i++) { tforce[3 * i] = hf[i]; tforce[3 * i + 1] = hf[nb + i]; tforce[3 * i + 2] = hf[2 * nb + i]; }
  }
  return SF_OK;
}

int sf_set_positions(sf_handle* h, const double* pos) {
  if (!h || !pos) { set_error("null argument"); return SF_E_ARG; }
  std::vector<double4> hp = tagged_positions(h, pos);
  HIPCHK(hipMemcpyAsync(h->d.pos, hp.data(), h->Nb * sizeof(double4), hipMemcpyHostToDevice, h->st));
  if (int rc = note_wrapped(h, pos)) return rc;
  if (int rc = launch_moved(h->d, h->st)) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return SF_OK;
}

int sf_get_state(sf_handle* h, double* pos, double* vel) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  const size_t nb = h->Nb;
  if (pos) {
    std::vector<double4> hp(nb);
    HIPCHK(hipMemcpyAsync(hp.data(), h->d.pos, nb * sizeof(double4), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t i = 0; i < nb; i++) { pos[3 * i] = hp[i].x; pos[3 * i + 1] = hp[i].y; pos[3 * i + 2] = hp[i].z; }
  }
  if (vel) {
    std::vector<double> hv(3 * nb);
    HIPCHK(hipMemcpyAsync(hv.data(), h->d.vx, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t i = 0; i < nb; i++) { vel[3 * i] = hv[i]; vel[3 * i + 1] = hv[nb + i]; vel[3 * i + 2] = hv[2 * nb + i]; }
  }
  return SF_OK;
}

int sf_get_thermostat(sf_handle* h, int replica, double* xi, double* eta) {
  if (!h || replica < 0 || replica >= h->R) { set_error("bad replica"); return SF_E_ARG; }
  std::vector<RepState> s;
  if (int rc = h->read_states(s)) return rc;
  int q = h->rp[replica].q_chain;
  for (int j = 0; j < q; j++) {
    if (xi) xi[j] = s[replica].xi[j];
    if (eta) eta[j] = s[replica].eta[j];
  }
  return q;
}

int sf_forces(sf_handle* h, double* f, double* fs, double* fb, double* fe, double* flj) {
  return sf_forces_ex(h, 1, f, fs, fb, fe, flj);
}

int sf_forces_ex(sf_handle* h, int force_rebuild, double* f, double* fs, double* fb, double* fe, double* flj) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  const size_t n3 = 3 * (size_t)h->Nb;
  if (!h->out_buf && h->alloc(&h->out_buf, 5 * n3)) return SF_E_HIP;
  for (int attempt = 0; attempt < 16; attempt++) {
    Dev d = h->d;  // capacities and buffers as regrown by the previous attempt
    d.out_f = h->out_buf; d.out_s = h->out_buf + n3; d.out_b = h->out_buf + 2 * n3;
    d.out_e = h->out_buf + 3 * n3; d.out_lj = h->out_buf + 4 * n3;
    if (int rc = launch_centres(d, h->st)) return rc;
    // a replay after regrowth always rebuilds (the regrown lists are empty)
    if (int rc = launch_rebuild(d, (force_rebuild || attempt > 0) ? 1 : 0, h->st)) return rc;
    if (int rc = h->launch_far(0)) return rc;
    if (int rc = launch_force(d, MODE_FORCES, 0, h->fp32, h->st)) return rc;
    int rc = h->check_and_fix();
    if (rc < 0) return rc;
    if (rc == 0) break;
    std::vector<RepState> s;
    if ((rc = h->read_states(s))) return rc;
    for (auto& x : s) { x.need_lj = x.need_q = x.need_cand = x.deformed = x.need_cell_lj = x.need_cell_q = 0; }
    HIPCHK(hipMemcpyAsync(h->d.rs, s.data(), h->R * sizeof(RepState), hipMemcpyHostToDevice, h->st));
  }
  if (int rc = launch_lists_done(h->d, h->st)) return rc;
  double* outs[5] = {f, fs, fb, fe, flj};
  for (int k = 0; k < 5; k++)
    if (outs[k]) HIPCHK(hipMemcpyAsync(outs[k], h->out_buf + k * n3, n3 * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SF_OK;
}

int sf_energies(sf_handle* h, double* totals, double* per_bead) {
  if (!h || !totals) { set_error("null argument"); return SF_E_ARG; }
  if (int rc = h->launch_far(1)) return rc;
  if (int rc = launch_energy(h->d, h->st)) return rc;
  std::vector<double> ep((size_t)h->Ns * 5);
  HIPCHK(hipMemcpyAsync(ep.data(), h->d.epart, ep.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  if (per_bead)
    HIPCHK(hipMemcpyAsync(per_bead, h->d.ebead, (size_t)h->Nb * 5 * sizeof(double), hipMemcpyDeviceToHost, h->st));
  std::vector<RepState> s;
  if (int rc = h->read_states(s)) return rc;
  for (int r = 0; r < h->R; r++) {
    const RepParams& P = h->rp[r];
    double se = 0, be = 0, ne = 0, ce = 0, ke = 0;
    for (int k = P.s0; k < P.s0 + P.ns; k++) {
      se += ep[5 * k]; be += ep[5 * k + 1]; ne += ep[5 * k + 2]; ce += ep[5 * k + 3]; ke += ep[5 * k + 4];
    }
    double tp = 0, tk = 0;
    for (int j = 0; j < P.q_chain; j++) {  // THERMOSTAT::potential_energy / kinetic_energy
      tp += P.dof[j] * 1.0 * P.T * s[r].eta[j];
      tk += 0.5 * P.Qb[j] * s[r].xi[j] * s[r].xi[j];
    }
    double* o = totals + 7 * r;
    o[0] = ke; o[1] = se; o[2] = be; o[3] = ne; o[4] = ce; o[5] = tp; o[6] = tk;
  }
  return SF_OK;
}

int sf_step(sf_handle* h, int n_steps) {
  if (!h || n_steps < 0) { set_error("bad arguments to sf_step"); return SF_E_ARG; }
  if (h->big_dirty) { set_error("big beads changed since the last sf_set_state: restart from the new set first"); return SF_E_ARG; }
  double ti = 0, trb = 0, tf = 0, tr = 0, nt = 0, tfar = 0;
  int done = 0, rc = SF_OK;
  int retries = 0;
  int goal = n_steps;
  // a replica that became non-finite stops the handle at its last finite state
  bool nonfinite = false;
  long long nf_step = 0;
  int nf_rep = -1;
  // a far-field model is audited every kChunk steps and timed steps launch directly, so both
  // keep the default length; otherwise one graph launch covers up to graph_steps steps
  const int chunk = (h->far_model || h->timing) ? kChunk : h->graph_steps;
  while (done < goal) {
    const int n = std::min(chunk, goal - done);
    if ((rc = h->snapshot(true))) break;
    if ((rc = h->launch_chunk(n))) break;
    rc = h->check_and_fix();
    if (rc < 0) break;
    if (rc == 2) {
      // back to the chunk's start, then forward to the step before the first non-finite one
      if (nonfinite) { set_error("non-finite state reappeared while replaying to the last finite step"); rc = SF_E_NONFINITE; break; }
      nonfinite = true;
      nf_step = h->bad_step;
      nf_rep = h->bad_rep;
      if ((rc = h->snapshot(false)) || (rc = h->clear_flags_force_rebuild())) break;
      std::vector<RepState> st0;
      if ((rc = h->read_states(st0))) break;
      const long long k = nf_step - 1 - st0[0].step;
      if (k < 0 || k >= n) { set_error("non-finite step outside the chunk that reported it"); rc = SF_E_NONFINITE; break; }
      goal = done + (int)k;
      rc = SF_OK;
      continue;
    }
    if (rc == 1) {
      if (++retries > 16) { set_error("list capacity did not converge"); rc = SF_E_OVERFLOW; break; }
      if ((rc = h->snapshot(false))) break;
      if ((rc = h->clear_flags_force_rebuild())) break;
      rc = SF_OK;
      continue;
    }
    if (h->far_model) {  // audit the far-field model after every chunk
      double err = 0;
      if ((rc = h->audit(&err))) break;
      if (!(err <= h->far_tol)) {
        // the chunk ran on a model that fails the audit: it is run again with the exact far field
        h->far_model = false;
        h->far_fell_back = true;
        if ((rc = h->snapshot(false)) || (rc = h->clear_flags_force_rebuild())) break;
        continue;
      }
    }
    if (h->timing) {
      for (int k = 0; k < n; k += kTimeEvery) {
        const hipEvent_t* e = &h->events[sf_handle::kEv * (k / kTimeEvery)];
        float a = 0, b = 0, fa = 0, c = 0, x = 0;
        (void)hipEventElapsedTime(&a, e[0], e[1]);
        (void)hipEventElapsedTime(&b, e[1], e[2]);
        (void)hipEventElapsedTime(&fa, e[2], e[3]);
        (void)hipEventElapsedTime(&c, e[3], e[4]);
        (void)hipEventElapsedTime(&x, e[4], e[5]);
        ti += a; trb += b; tfar += fa; tf += c; tr += x;
        nt += 1;
      }
    }
    done += n;
  }
  if (rc) return rc;
  if ((rc = launch_apply_com(h->d, h->st))) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  if (h->timing && nt > 0) {
    h->t_int = ti / nt; h->t_rb = trb / nt; h->t_force = tf / nt; h->t_red = tr / nt; h->t_steps = nt;
    h->t_far = tfar / nt;
  }
  if (nonfinite) {
    set_error("replica " + std::to_string(nf_rep) + " became non-finite at step " + std::to_string(nf_step) +
              "; the handle holds the last finite state (step " + std::to_string(nf_step - 1) + ")");
    return SF_E_NONFINITE;
  }
  return SF_OK;
}

int sf_set_timing(sf_handle* h, int enable) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  h->timing = enable != 0;
  return SF_OK;
}

int sf_get_stats(sf_handle* h, sf_stats* o) {
  if (!h || !o) { set_error("null argument"); return SF_E_ARG; }
  std::vector<RepState> s;
  if (int rc = h->read_states(s)) return rc;
  std::memset(o, 0, sizeof(*o));
  o->steps = s[0].step;
  for (auto& x : s) { o->rebuilds += x.rebuilds; o->lj_entries += x.lj_entries; o->q_entries += x.q_entries; }
  o->lj_capacity = h->d.lj_cap; o->q_capacity = h->d.q_cap;
  o->n_beads = h->Nb; o->n_subunits = h->Ns; o->n_charged = h->Nq; o->n_replicas = h->R;
  o->k_integrate_ms = h->t_int; o->k_rebuild_ms = h->t_rb; o->k_force_ms = h->t_force; o->k_reduce_ms = h->t_red;
  o->k_timed_steps = h->t_steps;
  o->graph_launches = h->graph_launches;
  o->graph_steps = h->graph_steps;
  o->k_far_ms = h->t_far;
  o->vcom = s[0].vcom;
  return SF_OK;
}

int sf_bd_uniforms(sf_handle* h, int replica, int64_t step, double* out) {
  if (!h || replica < 0 || replica >= h->R || step < 0 || !out) { set_error("bad arguments"); return SF_E_ARG; }
  const RepParams& P = h->rp[replica];
  // two engine calls per draw; the big beads' draws follow the subunit beads' every step
  uint64_t e = 6ull * ((uint64_t)P.nb + (uint64_t)h->d.nbig);
  // state before iteration `step`: seed * 16807^(6 (nb + n_big) step) mod (2^31-1)
  uint64_t st = P.seed;
  uint64_t mul = 1, base = lcg_pow(e);
  for (uint64_t k = (uint64_t)step; k; k >>= 1) {
    if (k & 1) mul = (mul * base) % kLcgMod;
    base = (base * base) % kLcgMod;
  }
  st = (st * mul) % kLcgMod;
  double* dout;
  if (h->alloc(&dout, 3 * (size_t)P.nb)) return SF_E_HIP;
  int rc = launch_bd_uniforms_state((uint32_t)st, P.nb, dout, h->d, h->st);
  if (!rc) {
    HIPCHK(hipMemcpyAsync(out, dout, 3 * (size_t)P.nb * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  h->release(dout);
  return rc;
}

int sf_set_graph_steps(sf_handle* h, int steps) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  if (steps > kMaxGraphSteps) { set_error("graph length above 20000 steps"); return SF_E_ARG; }
  h->graph_steps = steps <= 0 ? kChunk : steps;
  return SF_OK;
}

int sf_set_tile(sf_handle* h, int tile) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  if (!tile_supported(tile)) { set_error("tile must be 128, 256, 384 or 512"); return SF_E_ARG; }
  h->d.tile = tile;
  return SF_OK;
}

int sf_set_skin(sf_handle* h, int replica, double skin) {
  if (!h || replica < -1 || replica >= h->R) { set_error("bad replica"); return SF_E_ARG; }
  if (!(skin > 0)) { set_error("skin must be positive"); return SF_E_ARG; }
  const int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? h->R : replica + 1;
  for (int r = r0; r < r1; r++)
    if (skin > h->max_skin(r)) {
      set_error("box too small for this skin (minimum image needs cutoff + skin < box/2)");
      return SF_E_ARG;
    }
  for (int r = r0; r < r1; r++) h->put_skin(r, skin);
  h->fill_bounds();
  if (int rc = h->setup_cells()) return rc;
  if (int rc = h->upload_rp()) return rc;
  return h->request_rebuild();
}

int sf_get_skin(sf_handle* h, int replica, double* skin) {
  if (!h || replica < 0 || replica >= h->R || !skin) { set_error("bad arguments"); return SF_E_ARG; }
  *skin = h->prm[replica].skin;
  return SF_OK;
}

int sf_autotune(sf_handle* h, const sf_tune_opts* opt, sf_tune_result* res, double* trial_ms) {
  if (!h || !res) { set_error("null argument"); return SF_E_ARG; }
  static const double kSkins[] = {0.2, 0.3, 0.5, 0.8, 1.2};
  static const int kTiles[] = {128, 256, 384, 512};
  const double* skins = kSkins;
  const int* tiles = kTiles;
  int ns = (int)(sizeof(kSkins) / sizeof(kSkins[0])), nt = (int)(sizeof(kTiles) / sizeof(kTiles[0])), trial = 200;
  if (opt) {
    if (opt->n_skins > 0) { if (!opt->skins) { set_error("skins is null"); return SF_E_ARG; } skins = opt->skins; ns = opt->n_skins; }
    if (opt->n_tiles > 0) { if (!opt->tiles) { set_error("tiles is null"); return SF_E_ARG; } tiles = opt->tiles; nt = opt->n_tiles; }
    if (opt->trial_steps > 0) trial = opt->trial_steps;
  }
  for (int i = 0; i < ns; i++)
    if (!(skins[i] > 0)) { set_error("candidate skins must be positive"); return SF_E_ARG; }
  for (int j = 0; j < nt; j++)
    if (!tile_supported(tiles[j])) { set_error("candidate tiles must be 128, 256, 384 or 512"); return SF_E_ARG; }
  // save the dynamic state; every trial starts from it and the caller gets it back
  const size_t nb = h->Nb;
  double4* a_pos = nullptr; double* a_v = nullptr; double* a_f = nullptr; double* a_g = nullptr;
  RepState* a_rs = nullptr;
  if (h->alloc(&a_pos, nb) || h->alloc(&a_v, 3 * nb) || h->alloc(&a_f, 3 * nb) || h->alloc(&a_g, 3 * nb) ||
      h->alloc(&a_rs, h->R)) {
    h->release(a_pos); h->release(a_v); h->release(a_f); h->release(a_g); h->release(a_rs);
    return SF_E_HIP;
  }
  auto copy_state = [&](bool save) -> int {
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    Dev& d = h->d;
    HIPCHK(hipMemcpyAsync(save ? (void*)a_pos : (void*)d.pos, save ? (void*)d.pos : (void*)a_pos, nb * sizeof(double4), k, h->st));
    HIPCHK(hipMemcpyAsync(save ? a_v : d.vx, save ? d.vx : a_v, 3 * nb * sizeof(double), k, h->st));
    HIPCHK(hipMemcpyAsync(save ? a_f : d.fx, save ? d.fx : a_f, 3 * nb * sizeof(double), k, h->st));
    HIPCHK(hipMemcpyAsync(save ? a_g : d.gx, save ? d.gx : a_g, 3 * nb * sizeof(double), k, h->st));
    HIPCHK(hipMemcpyAsync(save ? (void*)a_rs : (void*)d.rs, save ? (void*)d.rs : (void*)a_rs, h->R * sizeof(RepState), k, h->st));
    if (!save)
      if (int rc = launch_centres(d, h->st)) return rc;
    HIPCHK(hipStreamSynchronize(h->st));
    return SF_OK;
  };
  // skin per replica: the candidate, or the replica's largest valid skin when its box is smaller
  auto apply = [&](double skin, int tile) -> int {
    for (int r = 0; r < h->R; r++) h->put_skin(r, std::min(skin, h->max_skin(r)));
    h->fill_bounds();
    h->d.tile = tile;
    if (int rc = h->setup_cells()) return rc;
    if (int rc = h->upload_rp()) return rc;
    return h->request_rebuild();
  };
  const bool timing = h->timing;
  const double t_save[6] = {h->t_int, h->t_rb, h->t_force, h->t_red, h->t_steps, h->t_far};
  h->timing = false;
  int rc = copy_state(true);
  double best = 1e300;
  int bi = 0, bj = 0, n = 0;
  for (int i = 0; i < ns && !rc; i++)
    for (int j = 0; j < nt && !rc; j++) {
      if ((rc = copy_state(false)) || (rc = apply(skins[i], tiles[j]))) break;
      // one chunk first: graph capture and list-capacity growth stay out of the timed steps
      if ((rc = sf_step(h, kChunk))) break;
      auto t0 = std::chrono::steady_clock::now();
      if ((rc = sf_step(h, trial))) break;
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial;
      if (trial_ms) trial_ms[i * nt + j] = ms;
      n++;
      if (ms < best) { best = ms; bi = i; bj = j; }
    }
  if (!rc) rc = copy_state(false);
  if (!rc) rc = apply(skins[bi], tiles[bj]);
  h->timing = timing;
  h->t_int = t_save[0]; h->t_rb = t_save[1]; h->t_force = t_save[2]; h->t_red = t_save[3]; h->t_steps = t_save[4];
  h->t_far = t_save[5];
  h->release(a_pos); h->release(a_v); h->release(a_f); h->release(a_g); h->release(a_rs);
  if (rc) return rc;
  res->skin = skins[bi];
  res->tile = tiles[bj];
  res->ms_per_step = best;
  res->n_trials = n;
  return SF_OK;
}

int sf_set_far_model(sf_handle* h, const sf_far_model* m, double tol) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  if (!h->d.split) { set_error("a far-field model needs the uncut Yukawa (ecut_el = 0)"); return SF_E_ARG; }
  if (!m) { h->far_model = false; return SF_OK; }
  if (h->mode != MODE_MD) { set_error("the far-field model is audited against MD forces; use MD"); return SF_E_ARG; }
  if (!(tol > 0)) { set_error("tolerance must be positive"); return SF_E_ARG; }
  FarSpec spec;
  const unsigned short* wsrc;
  const float* bsrc;
  far_model_device(m, &spec, &wsrc, &bsrc);
  if (spec.nq != h->t.nq()) { set_error("model and handle templates differ"); return SF_E_ARG; }
  for (int r = 0; r < h->R; r++) {
    if (h->far_rf[r] < spec.rf - 1e-9) { set_error("handle far_rf is below the model's split radius"); return SF_E_ARG; }
    const double k = h->prm[r].kappa;
    if (k < spec.kmin * (1 - 1e-9) || k > spec.kmax * (1 + 1e-9)) {
      set_error("a replica's salt is outside the model's training range"); return SF_E_ARG;
    }
  }
  if (!h->fdev.spec) {
    FarSpec* ds; unsigned short* dw; float* db; float* fr;
    if (h->alloc(&ds, 1) || h->alloc(&dw, kFPackN) || h->alloc(&db, 3 * kFHid + kFOutPad) ||
        h->alloc(&fr, (size_t)h->Ns * 12) || h->alloc(&h->qfar_exact, 3 * (size_t)std::max(h->Nq, 1)) ||
        h->alloc(&h->audit_out, 3 * (size_t)h->R))
      return SF_E_HIP;
    h->fdev.spec = ds; h->fdev.wpack = dw; h->fdev.bias = db; h->fdev.frames = fr;
  }
  HIPCHK(hipMemcpyAsync((void*)h->fdev.spec, &spec, sizeof(FarSpec), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync((void*)h->fdev.wpack, wsrc, kFPackN * sizeof(unsigned short), hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync((void*)h->fdev.bias, bsrc, (3 * kFHid + kFOutPad) * sizeof(float), hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->far_tol = tol;
  h->far_model = true;
  h->far_fell_back = false;
  h->far_audits = 0; h->far_last = 0; h->far_max = 0; h->far_last_force = 0; h->far_max_force = 0;
  return SF_OK;
}

int sf_far_drop(sf_handle* h, int drop) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  if (!h->d.split) { set_error("no far field: the Yukawa is cut"); return SF_E_ARG; }
  h->far_drop = drop != 0;
  return SF_OK;
}

int sf_far_audit(sf_handle* h, double* rel_err) {
  if (!h || !rel_err) { set_error("null argument"); return SF_E_ARG; }
  if (!h->far_model) { set_error("no far-field model in use"); return SF_E_ARG; }
  // the model's far field and the integrator force at the current positions
  if (int rc = h->launch_far(0)) return rc;
  return h->audit(rel_err);
}

int sf_get_far_stats(sf_handle* h, sf_far_stats* o) {
  if (!h || !o) { set_error("null argument"); return SF_E_ARG; }
  o->audits = h->far_audits; o->last_err = h->far_last; o->max_err = h->far_max;
  o->using_model = h->far_model ? 1 : 0; o->fell_back = h->far_fell_back ? 1 : 0;
  o->last_err_force = h->far_last_force; o->max_err_force = h->far_max_force;
  return SF_OK;
}

int sf_set_big_beads(sf_handle* h, int n, const double* xyz, double sigma, double charge, double shc) {
  if (!h || n < 0 || (n > 0 && !xyz)) { set_error("bad arguments to sf_set_big_beads"); return SF_E_ARG; }
  if (n > 0 && h->R != 1) { set_error("big beads need a single-replica handle"); return SF_E_ARG; }
  if (n > 0 && !(sigma > 0)) { set_error("big-bead diameter must be positive"); return SF_E_ARG; }
  HIPCHK(hipStreamSynchronize(h->st));  // no launch may read the tables being replaced
  h->drop_graph();
  RepParams& P = h->rp[0];
  if (n > 0) {
    const int np = h->t.np;
    if (n > h->nbig_cap) {
      if (h->alloc(&h->big_pos, (size_t)n)) return SF_E_HIP;
      h->nbig_cap = n;
    }
    if (!h->big_f && (h->alloc(&h->big_f, 2 * (size_t)h->Nb) || h->alloc(&h->big_tab, 2 * (size_t)np)))
      return SF_E_HIP;
    std::vector<double4> bp(n), tab(2 * (size_t)np);
    for (int i = 0; i < n; i++) bp[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
    // per template bead, as src/forces.cpp:209-219 forms them
    const double shc6 = std::pow(shc, 6);
    for (int l = 0; l < np; l++) {
      const double sig1 = sigma, sig2 = h->t.sigma[l], sigavg = (sig1 + sig2) / 2, delta = sigavg - shc;
      const double ys = std::exp(P.kappa * sigavg) / ((1 + P.kappa * sig1 / 2) * (1 + P.kappa * sig2 / 2));
      tab[2 * l] = make_double4(delta, delta + 1.12246205 * shc,
                                (delta + 1.12246205 * shc) * (delta + 1.12246205 * shc), ys);
      tab[2 * l + 1] = make_double4(charge * h->t.q[l] * P.lb, h->t.q[l] != 0 ? 1.0 : 0.0,
                                    P.ecut_el * P.ecut_el, shc6);
    }
    HIPCHK(hipMemcpy(h->big_pos, bp.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->big_tab, tab.data(), tab.size() * sizeof(double4), hipMemcpyHostToDevice));
  }
  h->d.nbig = n;
  h->d.big_pos = h->big_pos;
  h->d.big_tab = h->big_tab;
  h->d.big_f = h->big_f;
  // BD: each step draws the big beads' noise after the subunit beads' (src/md.cpp:482-487)
  P.mult6n = lcg_pow(6ull * ((unsigned long long)P.nb + (unsigned long long)n));
  HIPCHK(hipMemcpy(const_cast<RepParams*>(h->d.rp), &P, sizeof(RepParams), hipMemcpyHostToDevice));
  h->big_dirty = true;
  return SF_OK;
}

int sf_big_energies(sf_handle* h, double* out) {
  if (!h || !out) { set_error("null argument"); return SF_E_ARG; }
  out[0] = out[1] = 0;
  if (!h->d.nbig) return SF_OK;
  if (int rc = launch_big(h->d, h->st)) return rc;
  std::vector<double4> f(2 * (size_t)h->Nb);
  HIPCHK(hipMemcpyAsync(f.data(), h->big_f, f.size() * sizeof(double4), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  double ce = 0, ne = 0;
  for (size_t b = 0; b < (size_t)h->Nb; b++) { ne += f[2 * b].w; ce += f[2 * b + 1].w; }
  out[0] = ce; out[1] = ne;
  return SF_OK;
}

int sf_big_beads_place(const sf_template* t, int n_subunits, const double* pos, double box, int n, double sigma,
                       unsigned max_attempts, double* out) {
  if (!t || n_subunits < 0 || (n_subunits > 0 && !pos) || n < 0 || (n > 0 && !out) || !(box > 0)) {
    set_error("bad arguments to sf_big_beads_place");
    return SF_E_ARG;
  }
  // src/initialize.cpp:566-680 without a restart file; positions and box are long double there
  typedef long double LD;
  const LD bx = box;
  const int np = t->np;
  const size_t nb = (size_t)n_subunits * np;
  const double r2_big = sigma * sigma;
  auto wrap = [&](double v) { return v - bx * std::floor(v / bx + 0.5); };
  std::srand(0);
  for (int i = 0; i < n; i++) {
    bool placed = false;
    unsigned attempts = 0;
    while (!placed && attempts < max_attempts) {
      double c[3];
      for (int k = 0; k < 3; k++) c[k] = (double)((static_cast<double>(std::rand()) / RAND_MAX - 0.5) * bx);
      bool hit = false;
      for (size_t j = 0; j < nb && !hit; j++) {
        double dd[3];
        for (int k = 0; k < 3; k++) dd[k] = wrap(c[k] - (LD)pos[3 * j + k]);
        const double dist_sq = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        const double min_dist = t->sigma[j % np] / 2 + sigma / 2;
        hit = dist_sq < min_dist * min_dist;
      }
      for (int k = 0; k < i && !hit; k++) {
        double dd[3];
        for (int m = 0; m < 3; m++) dd[m] = wrap(c[m] - (LD)out[3 * k + m]);
        hit = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] < r2_big;
      }
      if (!hit) {
        for (int k = 0; k < 3; k++) out[3 * i + k] = c[k];
        placed = true;
      }
      attempts++;
    }
    if (attempts >= max_attempts) {
      set_error("failed to place big bead " + std::to_string(i + 1) + " after " + std::to_string(max_attempts) +
                " attempts");
      return SF_E_ARG;
    }
  }
  return SF_OK;
}

int sf_oligomers(sf_handle* h, double contact, int* label, int* mass_spectrum) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  if (!(contact > 0) || !std::isfinite(contact)) { set_error("contact distance must be positive"); return SF_E_ARG; }
  OligoBuf& o = h->olig;
  if (!o.parent) {
    if (h->alloc(&o.ext, h->R) || h->alloc(&o.parent, h->Ns) || h->alloc(&o.size, h->Ns) ||
        h->alloc(&o.label, h->Ns) || h->alloc(&o.spectrum, h->Ns) || h->alloc(&o.total, 1))
      return SF_E_HIP;
  }
  if (int rc = launch_centres(h->d, h->st)) return rc;
  if (int rc = launch_oligomers(h->d, contact, o, h->st)) return rc;
  std::vector<int> lab(h->Ns), ms(h->Ns);
  HIPCHK(hipMemcpyAsync(lab.data(), o.label, h->Ns * sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(ms.data(), o.spectrum, h->Ns * sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (label) std::memcpy(label, lab.data(), h->Ns * sizeof(int));
  if (mass_spectrum) std::memcpy(mass_spectrum, ms.data(), h->Ns * sizeof(int));
  int n = 0;
  for (int s = 0; s < h->Ns; s++) n += ms[s];
  return n;
}

int sf_oligomers_noalloc(sf_handle* h, double contact, int* label, int* mass_spectrum) {
  if (!h) { set_error("null handle"); return SF_E_ARG; }
  if (!(contact > 0) || !std::isfinite(contact)) { set_error("contact distance must be positive"); return SF_E_ARG; }
  OligoBuf o{};
  if (int rc = oligo_scratch(h->dev, h->Ns, h->R, o)) return rc;
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  if (cur != h->dev) HIPCHK(hipSetDevice(h->dev));
  int rc = launch_centres(h->d, h->st);
  if (!rc) rc = launch_oligomers(h->d, contact, o, h->st);
  int n = -1;
  if (!rc) {
    // straight into the caller's arrays: no staging buffer
    hipError_t e = hipSuccess;
    if (label) e = hipMemcpyAsync(label, o.label, h->Ns * sizeof(int), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess && mass_spectrum)
      e = hipMemcpyAsync(mass_spectrum, o.spectrum, h->Ns * sizeof(int), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipMemcpyAsync(&n, o.total, sizeof(int), hipMemcpyDeviceToHost, h->st);
    if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    if (e != hipSuccess) { set_error(std::string("sf_oligomers_noalloc: ") + hipGetErrorString(e)); rc = SF_E_HIP; }
  }
  if (cur != h->dev) (void)hipSetDevice(cur);
  return rc ? rc : n;
}

}  // extern "C"
