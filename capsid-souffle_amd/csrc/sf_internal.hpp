// Internal types shared by the host library and the HIP kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "souffle_gpu.h"

struct sf_template {
  // header values (src/initialize.cpp:47)
  double bondlength = 0, si_mass = 0, si_sigma = 0;
  int np = 0, ne = 0, nf = 0;
  // per template bead
  std::vector<double> x, y, z, q, sigma, mass;
  std::vector<int> type;
  // edges and faces, template order
  std::vector<int> e1, e2, etype;
  std::vector<double> len0;
  std::vector<int> f1, f2, f3;
  // ordered attractive (type_i, type_j) rows of lj_a
  std::vector<int> la1, la2;

  // ---- derived once per template
  int centre = 0;                 // bead closest to the template centroid
  double r_all = 0, r_q = 0;      // max distance from the centre bead: all / charged beads
  double diam = 0;                // largest bead-bead distance
  // body frame of a subunit: e1 along (x_fp1 - x_centre), e2 the part of (x_fp2 - x_centre)
  // normal to e1, e3 = e1 x e2; qbody = charged beads' template coordinates in that frame
  int fp1 = 0, fp2 = 0;
  bool frame_ok = false;          // fp2 is off the e1 axis
  std::vector<double> qbody;      // [nq][3]
  std::vector<int> hinges;        // [nh][4] B0, B1, across(F0), across(F1)
  std::vector<int> e_off, e_other;  // per-bead incident edges (CSR, ascending edge id)
  std::vector<double> e_len0;
  std::vector<int> h_off, h_slot;   // per-bead hinge incidences (CSR), slot = 4 h + k
  std::vector<int> qslot;           // charged index of each bead, -1 if uncharged
  std::vector<int> qlocal;          // template index of each charged bead
  std::vector<unsigned long long> att_mask;  // [64]: bit tj of att_mask[ti] = (ti,tj) attractive
  int nh() const { return (int)hinges.size() / 4; }
  int nq() const { return (int)qlocal.size(); }
};

namespace sf {

constexpr int kBlock = 128;      // threads of a subunit workgroup; templates need np <= kBlock
constexpr int kMaxChain = 8;     // Nose-Hoover chain length limit
constexpr int kMaxTypes = 64;    // bead type ids must be < kMaxTypes
constexpr int kMaxCand = 1024;   // candidate subunits per subunit at a list build
constexpr int kPartStride = 10;  // per-subunit partial sums (slot-major: part[k * Ns + s])
constexpr int kPartSlots = 9;    // slots in use: KE, sum v (3), BD noise (3), LJ and Yukawa list sizes
constexpr double kWcaFactor = 1.12246205;  // src/forces.cpp:27, src/energies.cpp:67
constexpr uint32_t kLcgMod = 2147483647u;  // minstd_rand0: x <- 16807 x mod (2^31 - 1)
constexpr uint32_t kLcgMul = 16807u;

void set_error(const std::string& s);

// Body frame from u1 = x_fp1 - x_centre and u2 = x_fp2 - x_centre: e1 = u1/|u1|,
// e2 = (u2 - (u2.e1) e1)/|...|, e3 = e1 x e2. R is row-major with columns e1 e2 e3, so
// lab = R body and body = R^T lab.
__host__ __device__ inline void body_frame(const double* u1, const double* u2, double* R) {
  const double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  const double e1[3] = {u1[0] / n1, u1[1] / n1, u1[2] / n1};
  const double a = u2[0] * e1[0] + u2[1] * e1[1] + u2[2] * e1[2];
  double w[3] = {u2[0] - a * e1[0], u2[1] - a * e1[1], u2[2] - a * e1[2]};
  const double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double e2[3] = {w[0] / nw, w[1] / nw, w[2] / nw};
  const double e3[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  for (int k = 0; k < 3; k++) { R[3 * k] = e1[k]; R[3 * k + 1] = e2[k]; R[3 * k + 2] = e3[k]; }
}

// Bead tags: subunit id in the high 32 bits, template-local index in the low 32 bits,
// carried bit-for-bit in the w component of the position double4.
inline double make_tag(int s, int l) {
  unsigned long long u = ((unsigned long long)(unsigned)s << 32) | (unsigned)l;
  double d;
  __builtin_memcpy(&d, &u, 8);
  return d;
}

constexpr int kMaxDeg = 8;        // edges per bead (padded rows)
constexpr int kMaxHinc = 16;      // hinge incidences per bead (padded rows)
constexpr int kRbRecords = 1536;  // LDS records of candidate beads at a list build
constexpr int kRbPipe = 8;        // candidates whose beads a list build fetches together
constexpr int kRbGrid = 8;        // local build grid: at most 8 cells per axis
constexpr int kRbLjBuf = 24;      // per-thread LDS buffer of a bead's LJ list at a build

// Device view of the template tables.
struct TopoDev {
  int np, nh, nq, centre;
  int fp1, fp2;                   // frame beads of the body frame
  const double* mass;
  const double* sigma;
  const int* type;
  const int* qslot;
  const int* qlocal;
  // per-bead padded rows: incident edges (other bead, rest length) and hinge incidences
  // (slot 4 h + k), so a thread fetches its whole row with independent loads
  const int* e_cnt;               // [np]
  const unsigned char* e_other;   // [np][kMaxDeg]
  const double* e_len0;           // [np][kMaxDeg]
  const int* h_cnt;               // [np]
  const unsigned short* h_slot;   // [np][kMaxHinc]
  const int* hinge;      // [nh][4]
  const unsigned long long* att_mask;  // [kMaxTypes]
};

// Per-replica constants.
struct RepParams {
  double box, hbox, ibox;        // ibox = 1 / box
  double ecut, ecut2, ecut_el, ecut_el2, kappa, lb, elj_att;
  double ks, kb, dt, damp, T;
  double skin, half_skin2;       // half_skin2 = (skin/2 - margin)^2, margin covering FP32 xb
  double sub_rc_lj2, sub_rc_q2;  // candidate-subunit centre distance^2 for a list build
  double rc_q_list2;             // (ecut_el + skin)^2
  double r_bound2;               // largest allowed bead distance^2 from its centre bead
  double r_bound;                // its square root
  double r_bound_q2;             // the same bound^2 for charged beads
  double rl_lj;                  // largest LJ/WCA list radius (cutoff + skin) over bead pairs
  double pair_reach;             // bound on the distance of any list entry (list radius + skin)
  int wrapped;                   // all positions lie in [-box/2, box/2] (kept so by the steps)
  int grid_ok;                   // 1 if the local build grid fits in half a box
  double inv_nb;                 // 1.0 / beads of the replica
  double Qb[kMaxChain];          // bath masses
  double inv_Qb[kMaxChain];      // 1 / Qb (0 where Qb = 0)
  double dof[kMaxChain];         // bath degrees of freedom (as double)
  int s0, ns, b0, nb, q0, nqr;   // subunit / bead / charged-slot ranges
  int q_chain, brownian, pair_fp32, split;  // split: uncut Yukawa as near (lists) + far field
  double far_rf2;                // far field: centre-bead distance^2 beyond this
  // cell grids of the list build (cut Yukawa): g cells per axis (1 or >= 3), first cell index
  // of the replica in the global cell arrays, 1 / cell edge
  int g_lj, g_q, c0_lj, c0_q;
  float inv_a_lj, inv_a_q;
  uint32_t seed, mult6n;         // engine seed; 16807^(6 nb) mod (2^31-1)
  const double* qq;              // [nq][nq] q_i q_j lb
  const double* yk;              // [nq][nq] Yukawa size factor exp(k shc)/((1+k s_i/2)(1+k s_j/2))
  const double* bd_amp;          // [np] sqrt(24 m / (damp dt))
};

// Per-replica dynamic state, device resident.
struct RepState {
  double xi[kMaxChain], eta[kMaxChain];      // after the last completed step
  double xi_n[kMaxChain], eta_n[kMaxChain];  // after the current step's first thermostat half
  double expfac;    // exp(-dt/2 xi_0) of the current step
  double kick_c;    // 0.5 dt sqrt(expfac)
  double ke;        // particle kinetic energy used by the chain
  double mean_v[3];     // centre-of-mass velocity still to be removed
  double mean_fran[3];  // mean BD noise of the current step
  double vcom;          // |mean v| removed at the last step
  uint32_t rng;         // engine state before the current step's draws
  int rebuild;          // list rebuild requested (set by the integrate kernel)
  int need_lj, need_q;  // longest list seen when it exceeded the capacity
  int need_cand;        // candidate subunits when they exceeded kMaxCand
  int deformed;         // a bead left the centre-bead bound at a list build
  int need_cell_lj, need_cell_q;  // largest cell occupancy seen when it exceeded the capacity
  int nonfinite;        // the kinetic energy or the centre-of-mass velocity became non-finite
  int resume;           // a (re)start takes ke from this state instead of measuring it
  long long bad_step;   // step counter value of the first non-finite step
  long long step, rebuilds, lj_entries, q_entries;
};

struct Dev {
  TopoDev topo;
  const RepParams* rp;
  RepState* rs;
  const int* sub_rep;
  double4* pos;        // [Nb] x, y, z, tag
  double* vx; double* vy; double* vz;
  double* fx; double* fy; double* fz;  // integrator force (tforce)
  double* gx; double* gy; double* gz;  // BD drag force
  float4* xb;          // positions at the last list build (FP32; the displacement test keeps a margin)
  double4* centres;    // [Ns] centre-bead positions
  int* lj_list; int* lj_cnt;
  int* q_list; int* q_cnt;
  double* part;        // [kPartStride][Ns]
  double* epart;       // [Ns][5] energy partial sums
  double* ebead;       // [Nb][5] per-bead energies
  double* out_f; double* out_s; double* out_b; double* out_e; double* out_lj;  // [Nb][3], nullable
  double4* qpos;       // [Nq] charged-bead positions (subunit-major, charged-slot order)
  double* qfar;        // [Nq][3] far-field Yukawa force on each charged bead (split mode)
  double* efar;        // [Nq] far-field Yukawa energy of each charged bead, pairs to higher ids
  // cell-list build: bead records (x, y, z, id) per cell, fixed capacity, occupancy counts
  float4* cell_lj; int* cnt_lj;
  float4* cell_q; int* cnt_q;
  int ccap_lj, ccap_q, ncell_lj, ncell_q;
  int cells;           // 1: lists come from the cell build (cut Yukawa)
  int lj_cap, q_cap;
  int Nb, Ns, Nq, R;
  int split;           // 1: every replica uses the uncut near/far Yukawa split
  int tile;            // list entries k_force stages in LDS per gather round (128, 256, 384, 512)
  // encapsulation big beads (single-replica handles): immobile positions, per-template-bead
  // constants [2 np] = (delta, wca_cut, wca_cut^2, yukawa_scale), (q_big q lb, charged,
  // ecut_el^2, shc^6), and, for the energy entries, the big-bead terms of each subunit bead
  // [2 Nb] = (ljforce, ne), (eforce, ce)
  int nbig;
  // BD noise: 16807^(6 b) mod (2^31 - 1) for replica-local bead b, the engine's skip-ahead to
  // the bead's first draw of a step (null for MD handles)
  const uint32_t* bd_skip;
  const double4* big_pos;
  const double4* big_tab;
  double4* big_f;
};

enum { MODE_FORCES = 0, MODE_MD = 1, MODE_BD = 2 };

// Device buffers of the oligomer analysis (analysis.hip): per subunit of the handle, and the
// per-replica bead extent.
struct OligoBuf {
  unsigned long long* ext;  // [R] largest squared bead distance from its centre bead (bits)
  int* parent;              // [Ns] union-find parent (global subunit index)
  int* size;                // [Ns] oligomer size at its root
  int* label;               // [Ns] replica-local index of the oligomer's smallest subunit
  int* spectrum;            // [Ns] per replica (at its first subunit): oligomers by size
  int* total;               // [1] oligomers over all replicas
};

// Largest handle (total subunits) the allocation-free oligomer entry point serves: its
// buffers live in a fixed device array of analysis.hip (6 MB + 256 B per GPU).
constexpr int kOligoScratchSubunits = 1 << 18;

// kernel launchers (kernels.hip)
int launch_centres(const Dev& d, hipStream_t st);
int launch_integrate(const Dev& d, int mode, hipStream_t st);
int launch_rebuild(const Dev& d, int force_rebuild, hipStream_t st);
int launch_moved(const Dev& d, hipStream_t st);       // after positions set from outside the loop
int launch_lists_done(const Dev& d, hipStream_t st);  // clears the pending-build flags
int launch_force(const Dev& d, int mode, int init, int fp32, hipStream_t st);  // uses d.tile
bool tile_supported(int tile);
int launch_finalize(const Dev& d, int init, hipStream_t st);
int launch_far_exact(const Dev& d, int energy, hipStream_t st);
int launch_fill(double* p, size_t n, double v, hipStream_t st);  // a kernel, so graphs may hold it
int launch_apply_com(const Dev& d, hipStream_t st);
int launch_energy(const Dev& d, hipStream_t st);
int launch_big(const Dev& d, hipStream_t st);
int launch_bd_uniforms_state(uint32_t state, int nb, double* out_dev, const Dev& d, hipStream_t st);
int launch_oligomers(const Dev& d, double contact, OligoBuf& o, hipStream_t st);  // analysis.hip
// Carves OligoBuf for ns subunits and R replicas out of device `dev`'s fixed scratch array
// (analysis.hip); the array's address is resolved once per device. No allocation.
int oligo_scratch(int dev, int ns, int R, OligoBuf& o);

// host helpers (topology.cpp)
uint32_t lcg_pow(uint64_t e);   // 16807^e mod (2^31-1)
bool derive_template(sf_template& t);

}  // namespace sf
