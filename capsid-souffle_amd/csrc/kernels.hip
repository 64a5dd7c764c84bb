// HIP kernels for gfx950 (MI355X). One SOUFFLE step is five launches:
//   k_integrate   one thread per bead: pending centre-of-mass removal, first half kick (NH) or
//                 drag + half kick (BD), drift, periodic wrap, list-displacement check, cell counts
//                 cleared for the next build
//   k_cell_fill   one 128-thread workgroup per subunit, returns at once unless a bead of its
//   k_cell_query  replica moved more than skin/2: cell lists of FP32 records, then each subunit's
//                 Verlet lists from the records of other subunits gathered into LDS (cut Yukawa;
//                 the uncut Yukawa's near lists come from k_rebuild)
//   k_force       one 128-thread workgroup per subunit: stretch + bending (hinge terms staged in
//                 LDS, gathered per bead through padded incidence rows) + LJ/WCA and Yukawa
//                 gathers over the lists + second half kick (NH) or noise + drag + half kick (BD)
//                 + per-subunit sums
//   k_finalize    one workgroup per replica: fixed-order reductions, Nose-Hoover chain,
//                 centre-of-mass velocity, BD engine advance
// No atomics touch forces or sums: each bead's force is gathered by one thread, its lists are
// in ascending bead id, and every reduction has a fixed shape, so runs are bit-reproducible and
// do not depend on when lists were rebuilt.
#include <hip/hip_runtime.h>

#include "sf_internal.hpp"

namespace sf {

__device__ __forceinline__ double min_img(double v, double box, double hbox) {
  // single conditional shift per axis, as src/forces.cpp:85-91 and src/functions.h:42-56, as a
  // select (v - (-box) is v + box exactly), so that a wave does not branch on it
  const double sh = v > hbox ? box : (v < -hbox ? -box : 0.0);
  return v - sh;
}
// The same single shift for a difference whose true separation is far below half a box (beads of
// one subunit, list neighbours, a bead and its position at the last list build): the shift is
// the nearest multiple of the box, formed without a branch. Identical to min_img for such
// differences (they are never within rounding of half a box).
__device__ __forceinline__ double min_img_near(double v, double box, double ibox) {
  return fma(-rint(v * ibox), box, v);
}
// min_img_near where the caller may know that no shift is needed: a difference of two raw
// coordinates less than half a box apart (two beads of a subunit that does not straddle the
// periodic boundary; a bead away from the box faces and a list neighbour), for which the shift
// rounds to zero and min_img_near returns the difference itself, bit for bit.
template <bool WRAP>
__device__ __forceinline__ double near_diff(double v, double box, double ibox) {
  return WRAP ? min_img_near(v, box, ibox) : v;
}
template <bool B>
struct Flag {
  static constexpr bool value = B;
};

// Stores of bulk per-bead data that the following kernels read (positions, velocities, forces,
// lists, cell records; not the positions of the last build, which every step reads back):
// nontemporal, so they stream to memory instead of waiting in this XCD's L2
// for the write-back at the kernel's end (the readers run on every XCD and read through memory
// anyway). The step is ~2% shorter with them at 2000 subunits.
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) { __builtin_nontemporal_store(v, p); }
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_stream(double4* p, double4 v) {
  dvec2* q = reinterpret_cast<dvec2*>(p);
  __builtin_nontemporal_store((dvec2){v.x, v.y}, q);
  __builtin_nontemporal_store((dvec2){v.z, v.w}, q + 1);
}
__device__ __forceinline__ void st_stream(float4* p, float4 v) {
  __builtin_nontemporal_store((fvec4){v.x, v.y, v.z, v.w}, reinterpret_cast<fvec4*>(p));
}
// Branch-free single shift in FP32 for the list builds: the same value as the conditional shift
// for a separation far below half a box; elsewhere both give a separation of about half a box,
// so a test against a list radius (< half a box) decides the same.
__device__ __forceinline__ float min_imgf_near(float v, float box, float ibox) { return fmaf(-rintf(v * ibox), box, v); }

// Template index of bead j (subunit-major numbering) by a float reciprocal and an exact fix-up.
__device__ __forceinline__ int local_of(int j, int np, float inv_np) {
  int q = (int)((float)j * inv_np);
  int r = j - q * np;
  while (r < 0) r += np;
  while (r >= np) r -= np;
  return r;
}

// The same for j < 2^23 (the cell builds' range): the float quotient is then within 1/np of
// j / np, so one correction step each way is exact (no loop).
__device__ __forceinline__ int local_of_small(int j, int np, float inv_np) {
  const int q = (int)((float)j * inv_np);
  int r = j - q * np;
  r = r < 0 ? r + np : r;
  return r >= np ? r - np : r;
}

__device__ __forceinline__ int tag_sub(double w) { return (int)(__double_as_longlong(w) >> 32); }
__device__ __forceinline__ int tag_loc(double w) { return (int)(__double_as_longlong(w) & 0xffffffffll); }

// x * y mod (2^31 - 1) for x, y < 2^31 (Mersenne reduction).
__device__ __forceinline__ uint32_t lcg_mulmod(uint32_t x, uint32_t y) {
  unsigned long long p = (unsigned long long)x * y;
  unsigned long long r = (p & 0x7fffffffull) + (p >> 31);
  if (r >= kLcgMod) r -= kLcgMod;
  return (uint32_t)r;
}
__device__ uint32_t lcg_powmod(unsigned long long e) {
  uint32_t result = 1, base = kLcgMul;
  while (e) {
    if (e & 1) result = lcg_mulmod(result, base);
    base = lcg_mulmod(base, base);
    e >>= 1;
  }
  return result;
}

// One uniform(-0.5, 0.5) value of std::uniform_real_distribution<double> over
// std::minstd_rand0: generate_canonical<double, 53> takes two engine outputs g1, g2 and
// returns ((g1 - 1) + (g2 - 1) * R) / R^2 with R = 2^31 - 2, each operation rounded in double
// (libstdc++ bits/random.tcc generate_canonical); the distribution then returns u * 1 + (-0.5).
__device__ __forceinline__ double canonical_uniform(uint32_t g1, uint32_t g2) {
#pragma clang fp contract(off)
  const double R = 2147483646.0;
  double sum = 0.0;
  sum += (double)(g1 - 1u) * 1.0;
  sum += (double)(g2 - 1u) * R;
  double tmp = R * R;
  double ret = sum / tmp;
  if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;
  return ret * 1.0 + (-0.5);
}

// The 3 uniform draws of a bead for the step whose engine state is `state`: draws 6b+1 .. 6b+6
// of the stream (b replica-local), x, y, z order (src/md.cpp:464-466); skip = 16807^(6b) mod
// (2^31 - 1).
__device__ __forceinline__ void bd_draws_skip(uint32_t state, uint32_t skip, double& ux, double& uy, double& uz) {
  uint32_t x = lcg_mulmod(state, skip);
  uint32_t g[6];
#pragma unroll
  for (int k = 0; k < 6; k++) { x = lcg_mulmod(x, kLcgMul); g[k] = x; }
  ux = canonical_uniform(g[0], g[1]);
  uy = canonical_uniform(g[2], g[3]);
  uz = canonical_uniform(g[4], g[5]);
}
__device__ __forceinline__ void bd_draws(uint32_t state, int b, double& ux, double& uy, double& uz) {
  bd_draws_skip(state, lcg_powmod(6ull * (unsigned long long)b), ux, uy, uz);
}

// Replica of subunit s; a handle holding one system skips the table lookup, so that the
// replica's parameters are requested without waiting for it.
__device__ __forceinline__ int rep_of(const Dev& d, int s) { return d.R == 1 ? 0 : d.sub_rep[s]; }

// Row shift of a double across lanes (DPP row_shr:N within rows of 16 lanes; lanes without a
// source get 0).
template <int N>
__device__ __forceinline__ double dpp_row_shr(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int l2 = __builtin_amdgcn_update_dpp(0, lo, 0x110 + N, 0xf, 0xf, true);
  const int h2 = __builtin_amdgcn_update_dpp(0, hi, 0x110 + N, 0xf, 0xf, true);
  return __hiloint2double(h2, l2);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Sum over the 64 lanes of a wave (all lanes active), valid in every lane as a uniform value:
// running sums within each row of 16 lanes by DPP shifts (the row total lands in its last lane),
// then the four row totals in order. A fixed shape, so the result is deterministic.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_row_shr<1>(v);
  v += dpp_row_shr<2>(v);
  v += dpp_row_shr<4>(v);
  v += dpp_row_shr<8>(v);
  return ((lane_value(v, 15) + lane_value(v, 31)) + lane_value(v, 47)) + lane_value(v, 63);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Nose-Hoover chain variable j (src/functions.cpp:33-58).
// invQ[j] = 1 / Qb[j], formed on the host (the same correctly rounded value as the division).
__device__ void chain_xi(int j, double* xi, const double* Qb, const double* invQ, const double* dof, double T,
                         double dt, double ke) {
  if (Qb[j] == 0) return;
  if (j != 0) {
    xi[j] = xi[j] * exp(-0.5 * dt * xi[j + 1]) +
            0.5 * dt * invQ[j] * (Qb[j - 1] * xi[j - 1] * xi[j - 1] - dof[j] * 1.0 * T) *
                exp(-0.25 * dt * xi[j + 1]);
  } else {
    xi[j] = xi[j] * exp(-0.5 * dt * xi[j + 1]) +
            0.5 * dt * invQ[j] * (2 * ke - dof[j] * 1.0 * T) * exp(-0.25 * dt * xi[j + 1]);
  }
}

// ------------------------------------------------------------------------------------------
__global__ void k_centres(Dev d) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.Ns) return;
  const int np = d.topo.np, nq = d.topo.nq;
  d.centres[s] = d.pos[(size_t)s * np + d.topo.centre];
  for (int k = 0; k < nq; k++) d.qpos[(size_t)s * nq + k] = d.pos[(size_t)s * np + d.topo.qlocal[k]];
}

// First half of the step (src/md.cpp:400-431): NH: v = v e + F (dt/2 sqrt(e))/m;
// BD: fdrag = -m v / damp, v += F dt/(2m). Then x += v dt and one periodic wrap
// (src/bead.h:72-81). The centre-of-mass velocity left by the previous step is removed first.
template <int MODE>
__global__ __launch_bounds__(256) void k_integrate(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  // subunit and template index from the bead index (beads are subunit-major), so the replica
  // lookup does not wait for the position load
  const int np = d.topo.np;
  const int l = local_of(b, np, 1.0f / (float)np);
  const int s = (int)((float)(b - l) * (1.0f / (float)np) + 0.5f);
  const int r = rep_of(d, s);
  double4 p = d.pos[b];
  const RepParams& P = d.rp[r];
  RepState& S = d.rs[r];
  const double box = P.box, hbox = P.hbox, dt = P.dt, half_skin2 = P.half_skin2;
  const double m = d.topo.mass[l];
  const double mvx = S.mean_v[0], mvy = S.mean_v[1], mvz = S.mean_v[2];
  const float4 q = d.xb[b];
  double vx = d.vx[b] - mvx, vy = d.vy[b] - mvy, vz = d.vz[b] - mvz;
  const double fx = d.fx[b], fy = d.fy[b], fz = d.fz[b];
  if (MODE == MODE_MD) {
    const double e = S.expfac, c = S.kick_c / m;
    vx = vx * e + fx * c; vy = vy * e + fy * c; vz = vz * e + fz * c;
  } else {
    const double cd = -1.0 * m / P.damp;
    st_stream(&d.gx[b], vx * cd); st_stream(&d.gy[b], vy * cd); st_stream(&d.gz[b], vz * cd);
    const double c = 0.5 * dt / m;
    vx = vx + fx * c; vy = vy + fy * c; vz = vz + fz * c;
  }
  p.x = p.x + vx * dt; p.y = p.y + vy * dt; p.z = p.z + vz * dt;
  p.x = min_img(p.x, box, hbox); p.y = min_img(p.y, box, hbox); p.z = min_img(p.z, box, hbox);  // one wrap
  st_stream(&d.pos[b], p);
  st_stream(&d.vx[b], vx); st_stream(&d.vy[b], vy); st_stream(&d.vz[b], vz);
  const double ibox = P.ibox;
  const double ddx = min_img_near(p.x - (double)q.x, box, ibox), ddy = min_img_near(p.y - (double)q.y, box, ibox),
               ddz = min_img_near(p.z - (double)q.z, box, ibox);
  if (ddx * ddx + ddy * ddy + ddz * ddz > half_skin2) S.rebuild = 1;  // same value from any thread
  if (l == d.topo.centre) d.centres[s] = p;
  const int qs = d.topo.qslot[l];
  if (qs >= 0) d.qpos[(size_t)s * d.topo.nq + qs] = p;
  // cell occupancies of this step's (possible) list build start from zero
  if (d.cells) {
    for (int c = b; c < d.ncell_lj; c += d.Nb) d.cnt_lj[c] = 0;
    for (int c = b; c < d.ncell_q; c += d.Nb) d.cnt_q[c] = 0;
  }
}

// Positions set from outside the step loop (sf_set_positions): centre beads and charged-bead
// copies refreshed, the replica's list build requested when a bead moved more than skin/2 from
// its position at the last build (the test of k_integrate), cell counts cleared for that build.
__global__ __launch_bounds__(256) void k_moved(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  const int np = d.topo.np;
  const int s = b / np, l = b - s * np;
  const int r = rep_of(d, s);
  const RepParams& P = d.rp[r];
  const double4 p = d.pos[b];
  const float4 q = d.xb[b];
  const double ddx = min_img(p.x - (double)q.x, P.box, P.hbox), ddy = min_img(p.y - (double)q.y, P.box, P.hbox),
               ddz = min_img(p.z - (double)q.z, P.box, P.hbox);
  if (!(ddx * ddx + ddy * ddy + ddz * ddz <= P.half_skin2)) d.rs[r].rebuild = 1;  // same value from any thread
  if (l == d.topo.centre) d.centres[s] = p;
  const int qs = d.topo.qslot[l];
  if (qs >= 0) d.qpos[(size_t)s * d.topo.nq + qs] = p;
  if (d.cells) {
    for (int c = b; c < d.ncell_lj; c += d.Nb) d.cnt_lj[c] = 0;
    for (int c = b; c < d.ncell_q; c += d.Nb) d.cnt_q[c] = 0;
  }
}

// Lists are current after a force evaluation outside the step loop: no build pending.
__global__ void k_lists_done(Dev d) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < d.R) d.rs[r].rebuild = 0;
}

// ------------------------------------------------------------------------------------------
// Neighbour lists.
// LJ list of bead i: beads j of other subunits within (cutoff of the pair + skin), the cutoff
// being ecut for an attractive (type_i, type_j) row of lj_a and 1.12246205 (s_i + s_j)/2
// otherwise (src/forces.cpp:112-138). Yukawa list of a charged bead: charged beads j != i of
// any subunit (src/forces.cpp:79) within ecut_el + skin. Both lists hold ascending global
// bead ids, the order in which the reference's update_pairlist appends
// (src/functions.cpp:93-110). Candidate subunits are those whose centre beads lie within
// the pair reach plus twice the centre-to-bead bound.

// Candidate subunits of s in ascending id: each thread scans a contiguous block of subunit
// ids so that one workgroup prefix sum places all candidates in order.
__device__ int find_candidates(const Dev& d, const RepParams& P, RepState& S, int s, int tid, double4 cs,
                               int* cand, unsigned char* cflag, int* scan) {
  const double box = P.box, hbox = P.hbox;
  const int nq = d.topo.nq;
  int total = 0;
  for (int base = 0; base < P.ns; base += 64 * kBlock) {
    const int span = min(P.ns - base, 64 * kBlock);
    const int per = (span + kBlock - 1) / kBlock;
    unsigned long long bl = 0, bq = 0;
    const int t0 = P.s0 + base + tid * per;
    const int lim = P.s0 + base + span;
    for (int m0 = 0; m0 < per; m0 += 8) {
      double4 ct[8];
#pragma unroll
      for (int u = 0; u < 8; u++)  // eight centres requested together
        if (m0 + u < per && t0 + m0 + u < lim) ct[u] = d.centres[t0 + m0 + u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int m = m0 + u, t = t0 + m;
        if (m >= per || t >= lim) break;
        const double dx = min_img(cs.x - ct[u].x, box, hbox), dy = min_img(cs.y - ct[u].y, box, hbox),
                     dz = min_img(cs.z - ct[u].z, box, hbox);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (t != s && d2 < P.sub_rc_lj2) bl |= 1ull << m;
        if (nq > 0 && d2 < P.sub_rc_q2) bq |= 1ull << m;
      }
    }
    const int cnt = __popcll(bl | bq);
    scan[tid] = cnt;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {  // inclusive Hillis-Steele scan
      const int v = (tid >= o) ? scan[tid - o] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    int slot = total + scan[tid] - cnt;
    for (int m = 0; m < per; m++) {
      const unsigned long long bit = 1ull << m;
      if ((bl | bq) & bit) {
        if (slot < kMaxCand) {
          cand[slot] = t0 + m;
          cflag[slot] = (unsigned char)(((bl & bit) ? 1 : 0) | ((bq & bit) ? 2 : 0));
        }
        slot++;
      }
    }
    total += scan[kBlock - 1];
    __syncthreads();
  }
  if (total > kMaxCand) {
    if (tid == 0) atomicMax(&S.need_cand, total);
    total = kMaxCand;
  }
  return total;
}

// Exact FP64 build over the candidates, one staged candidate at a time. Used when a subunit's
// neighbourhood overflows the LDS records of the fast build.
__device__ void build_lists_exact(const Dev& d, const RepParams& P, int s, int tid, double4 pi,
                                  const int* cand, const unsigned char* cflag, int nc, double4* stage,
                                  bool do_lj, bool do_q, int& cnt_lj, int& cnt_q) {
  const int np = d.topo.np, nq = d.topo.nq;
  const double box = P.box, hbox = P.hbox;
  const bool active = tid < np;
  const int gi = s * np + tid;
  const int ti = active ? d.topo.type[tid] : 0;
  const double si = active ? d.topo.sigma[tid] : 0.0;
  const unsigned long long mask_i = active ? d.topo.att_mask[ti] : 0ull;
  const int qi = active ? d.topo.qslot[tid] : -1;
  const int gq = qi >= 0 ? s * nq + qi : -1;
  const double ecut_l = P.ecut + P.skin;
  const double ecut_l2 = ecut_l * ecut_l;
  if (do_lj) cnt_lj = 0;
  if (do_q) cnt_q = 0;
  for (int c = 0; c < nc; c++) {
    const int t = cand[c];
    const int f = cflag[c];
    __syncthreads();
    if (tid < np) stage[tid] = d.pos[(size_t)t * np + tid];
    __syncthreads();
    if (!active) continue;
    if (do_lj && (f & 1)) {
      for (int k = 0; k < np; k++) {
        const double4 pj = stage[k];
        const double dx = min_img(pi.x - pj.x, box, hbox), dy = min_img(pi.y - pj.y, box, hbox),
                     dz = min_img(pi.z - pj.z, box, hbox);
        const double r2 = dx * dx + dy * dy + dz * dz;
        const int tj = d.topo.type[k];
        double rl2;
        if ((mask_i >> tj) & 1ull) {
          rl2 = ecut_l2;
        } else {
          const double w = kWcaFactor * 0.5 * (si + d.topo.sigma[k]) + P.skin;
          rl2 = w * w;
        }
        if (r2 < rl2) {
          if (cnt_lj < d.lj_cap) d.lj_list[(size_t)cnt_lj * d.Nb + gi] = t * np + k;
          cnt_lj++;
        }
      }
    }
    if (do_q && (f & 2) && qi >= 0) {
      for (int kq = 0; kq < nq; kq++) {
        const int k = d.topo.qlocal[kq];
        const int j = t * np + k;
        if (j == gi) continue;
        const double4 pj = stage[k];
        const double dx = min_img(pi.x - pj.x, box, hbox), dy = min_img(pi.y - pj.y, box, hbox),
                     dz = min_img(pi.z - pj.z, box, hbox);
        if (dx * dx + dy * dy + dz * dz < P.rc_q_list2) {
          if (cnt_q < d.q_cap) d.q_list[(size_t)cnt_q * d.Nq + gq] = j;
          cnt_q++;
        }
      }
    }
  }
}

// j mod np and j / np for j < 2^23 via a float reciprocal and one correction.

// Fast build. Candidate beads are taken relative to the centre bead of s (minimum image,
// then FP32). LJ candidates inside the cube |rel| <= r_bound + rl_lj are counted and placed
// into an LDS grid of at most 8^3 cells of edge >= rl_lj; each bead of s then tests the 27
// cells around it and keeps hits in a per-thread LDS buffer sorted by id. Yukawa candidates'
// charged beads are staged in candidate order and scanned in order by the charged beads.
// FP32 tests use the list radius plus 1e-3 sigma, so the lists are supersets of the exact
// FP64 ones; entries beyond a force cutoff contribute nothing.
__global__ __launch_bounds__(kBlock) void k_rebuild(Dev d, int force_rebuild) {
  __shared__ int cand[kMaxCand];
  __shared__ unsigned char cflag[kMaxCand];
  __shared__ int scan[kBlock];
  __shared__ float4 rec[kRbRecords];
  __shared__ int cell_cnt[kRbGrid * kRbGrid * kRbGrid];
  __shared__ int cell_off[kRbGrid * kRbGrid * kRbGrid + 1];
  __shared__ int ljbuf[kRbLjBuf * kBlock];  // also the Yukawa record masks (after the LJ build)
  __shared__ float4 qown[32];
  __shared__ int fallback_sh;
  __shared__ float tsig[kBlock];
  __shared__ int ttype[kBlock];
  double4* stage = reinterpret_cast<double4*>(rec);  // exact fallback and final sums reuse the records

  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = rep_of(d, s);
  RepState& S = d.rs[r];
  if (!force_rebuild && S.rebuild == 0) return;
  const RepParams& P = d.rp[r];
  const int np = d.topo.np, nq = d.topo.nq;
  const double box = P.box, hbox = P.hbox;
  const float boxf = (float)box, iboxf = (float)P.ibox;
  const double4 cs = d.centres[s];
  const bool active = tid < np;
  const int gi = s * np + tid;
  const double4 pi = active ? d.pos[gi] : make_double4(0, 0, 0, 0);
  if (tid == 0) fallback_sh = P.grid_ok ? 0 : 1;

  const int nc = find_candidates(d, P, S, s, tid, cs, cand, cflag, scan);

  // ---- LJ: local cell grid
  const float H = (float)(P.r_bound + P.rl_lj) + 1e-3f;
  int G = (int)floorf(2.0f * H / ((float)P.rl_lj + 1e-3f));
  if (G < 1) G = 1;
  if (G > kRbGrid) G = kRbGrid;
  const float a = 2.0f * H / (float)G;
  const float inv_a = 1.0f / a;
  const int ncell = G * G * G;
  for (int c = tid; c < ncell; c += kBlock) cell_cnt[c] = 0;
  if (active) { tsig[tid] = (float)d.topo.sigma[tid]; ttype[tid] = d.topo.type[tid]; }
  __syncthreads();
  // pass 1: count (the beads of kPipe candidates are fetched together)
  for (int c0 = 0; c0 < nc; c0 += kRbPipe) {
    double4 pj[kRbPipe];
    bool use[kRbPipe];
#pragma unroll
    for (int u = 0; u < kRbPipe; u++) {
      use[u] = active && c0 + u < nc && (cflag[c0 + u] & 1);
      if (use[u]) pj[u] = d.pos[(size_t)cand[c0 + u] * np + tid];
    }
#pragma unroll
    for (int u = 0; u < kRbPipe; u++) {
      if (!use[u]) continue;
      const float rx = (float)min_img(pj[u].x - cs.x, box, hbox), ry = (float)min_img(pj[u].y - cs.y, box, hbox),
                  rz = (float)min_img(pj[u].z - cs.z, box, hbox);
      if (fabsf(rx) <= H && fabsf(ry) <= H && fabsf(rz) <= H) {
        const int cx = min(G - 1, (int)((rx + H) * inv_a)), cy = min(G - 1, (int)((ry + H) * inv_a)),
                  cz = min(G - 1, (int)((rz + H) * inv_a));
        atomicAdd(&cell_cnt[(cz * G + cy) * G + cx], 1);
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int c = 0; c < ncell; c++) { cell_off[c] = acc; acc += cell_cnt[c]; }
    cell_off[ncell] = acc;
    if (acc > kRbRecords) fallback_sh = 1;
  }
  __syncthreads();
  for (int c = tid; c < ncell; c += kBlock) cell_cnt[c] = 0;
  __syncthreads();
  const bool fast = fallback_sh == 0;
  // pass 2: place
  if (fast) {
    for (int c0 = 0; c0 < nc; c0 += kRbPipe) {
      double4 pj[kRbPipe];
      bool use[kRbPipe];
#pragma unroll
      for (int u = 0; u < kRbPipe; u++) {
        use[u] = active && c0 + u < nc && (cflag[c0 + u] & 1);
        if (use[u]) pj[u] = d.pos[(size_t)cand[c0 + u] * np + tid];
      }
#pragma unroll
      for (int u = 0; u < kRbPipe; u++) {
        if (!use[u]) continue;
        const int j = cand[c0 + u] * np + tid;
        const float rx = (float)min_img(pj[u].x - cs.x, box, hbox), ry = (float)min_img(pj[u].y - cs.y, box, hbox),
                    rz = (float)min_img(pj[u].z - cs.z, box, hbox);
        if (fabsf(rx) <= H && fabsf(ry) <= H && fabsf(rz) <= H) {
          const int cx = min(G - 1, (int)((rx + H) * inv_a)), cy = min(G - 1, (int)((ry + H) * inv_a)),
                    cz = min(G - 1, (int)((rz + H) * inv_a));
          const int cell = (cz * G + cy) * G + cx;
          const int slot = cell_off[cell] + atomicAdd(&cell_cnt[cell], 1);
          rec[slot] = make_float4(rx, ry, rz, __int_as_float(j));
        }
      }
    }
  }
  __syncthreads();
  int cnt_lj = 0, cnt_q = 0;
  if (fast && active) {
    const int ti = d.topo.type[tid];
    const float si = (float)d.topo.sigma[tid];
    const unsigned long long mask_i = d.topo.att_mask[ti];
    const float ratt = (float)(P.ecut + P.skin) + 1e-3f;
    const float ratt2 = ratt * ratt;
    const float skinf = (float)P.skin + 1e-3f;
    const float inv_np = 1.0f / (float)np;
    const float ix = (float)min_img(pi.x - cs.x, box, hbox), iy = (float)min_img(pi.y - cs.y, box, hbox),
                iz = (float)min_img(pi.z - cs.z, box, hbox);
    const int cx = min(G - 1, max(0, (int)((ix + H) * inv_a))), cy = min(G - 1, max(0, (int)((iy + H) * inv_a))),
              cz = min(G - 1, max(0, (int)((iz + H) * inv_a)));
    bool over = false;
    for (int zz = max(0, cz - 1); zz <= min(G - 1, cz + 1); zz++)
      for (int yy = max(0, cy - 1); yy <= min(G - 1, cy + 1); yy++)
        for (int xx = max(0, cx - 1); xx <= min(G - 1, cx + 1); xx++) {
          const int cell = (zz * G + yy) * G + xx;
          for (int m = cell_off[cell]; m < cell_off[cell + 1]; m++) {
            const float4 rj = rec[m];
            const int j = __float_as_int(rj.w);
            const int lj = local_of(j, np, inv_np);
            const int tj = ttype[lj];
            const float dx = min_imgf_near(ix - rj.x, boxf, iboxf), dy = min_imgf_near(iy - rj.y, boxf, iboxf),
                        dz = min_imgf_near(iz - rj.z, boxf, iboxf);
            const float r2 = dx * dx + dy * dy + dz * dz;
            float thr2;
            if ((mask_i >> tj) & 1ull) {
              thr2 = ratt2;
            } else {
              const float w = (float)kWcaFactor * 0.5f * (si + tsig[lj]) + skinf;
              thr2 = w * w;
            }
            if (r2 < thr2) {
              if (cnt_lj < kRbLjBuf) {
                int p = cnt_lj;
                while (p > 0 && ljbuf[(p - 1) * kBlock + tid] > j) {
                  ljbuf[p * kBlock + tid] = ljbuf[(p - 1) * kBlock + tid];
                  p--;
                }
                ljbuf[p * kBlock + tid] = j;
              } else {
                over = true;
              }
              cnt_lj++;
            }
          }
        }
    if (over) fallback_sh = 1;
  }
  __syncthreads();
  const bool lj_fast = fallback_sh == 0;
  if (lj_fast && active) {
    for (int k = 0; k < cnt_lj && k < d.lj_cap; k++) st_stream(&d.lj_list[(size_t)k * d.Nb + gi], ljbuf[k * kBlock + tid]);
  }
  // ---- Yukawa: charged beads of the candidates, in candidate order
  bool q_fast = true;
  if (nq > 0) {
    int nqc = 0;  // Yukawa candidates are a prefix-ordered subset of cand
    __syncthreads();
    if (tid == 0) {
      int k = 0;
      for (int c = 0; c < nc; c++)
        if (cflag[c] & 2) {
          if (k < kRbGrid * kRbGrid * kRbGrid) cell_cnt[k] = cand[c];  // reuse cell_cnt as the ordered list
          k++;
        }
      cell_off[0] = k;
    }
    __syncthreads();
    nqc = cell_off[0];
    const int nrec = nqc * nq;
    q_fast = nrec <= kRbRecords && nqc <= kRbGrid * kRbGrid * kRbGrid && nq <= 32;
    if (q_fast) {
      // stage the candidates' charged beads in candidate order, and this subunit's own
      const int qi = active ? d.topo.qslot[tid] : -1;
      if (qi >= 0)
        qown[qi] = make_float4((float)min_img(pi.x - cs.x, box, hbox), (float)min_img(pi.y - cs.y, box, hbox),
                               (float)min_img(pi.z - cs.z, box, hbox), __int_as_float(gi));
      for (int f = tid; f < nrec; f += kBlock) {
        const int c = f / nq, kq = f - c * nq;
        const int j = cell_cnt[c] * np + d.topo.qlocal[kq];
        const double4 pj = d.pos[j];
        rec[f] = make_float4((float)min_img(pj.x - cs.x, box, hbox), (float)min_img(pj.y - cs.y, box, hbox),
                             (float)min_img(pj.z - cs.z, box, hbox), __int_as_float(j));
      }
      __syncthreads();
      // every thread tests its records against all charged beads of s: bit kq of ymask[f] says
      // record f is in the list of charged bead kq
      unsigned* ymask = reinterpret_cast<unsigned*>(ljbuf);
      const float rq = sqrtf((float)P.rc_q_list2) + 1e-3f;
      const float rq2 = rq * rq;
      for (int f = tid; f < nrec; f += kBlock) {
        const float4 rj = rec[f];
        const int j = __float_as_int(rj.w);
        unsigned msk = 0;
        for (int kq = 0; kq < nq; kq++) {
          const float4 oi = qown[kq];
          const float dx = min_imgf_near(oi.x - rj.x, boxf, iboxf), dy = min_imgf_near(oi.y - rj.y, boxf, iboxf),
                      dz = min_imgf_near(oi.z - rj.z, boxf, iboxf);
          if (j != __float_as_int(oi.w) && dx * dx + dy * dy + dz * dz < rq2) msk |= 1u << kq;
        }
        ymask[f] = msk;
      }
      __syncthreads();
      if (qi >= 0) {
        const int gq = s * nq + qi;
        for (int f = 0; f < nrec; f++) {
          if ((ymask[f] >> qi) & 1u) {
            if (cnt_q < d.q_cap) st_stream(&d.q_list[(size_t)cnt_q * d.Nq + gq], __float_as_int(rec[f].w));
            cnt_q++;
          }
        }
      }
    }
  }
  // ---- exact fallback for whatever overflowed the fast path (uniform branches)
  if (!lj_fast || !q_fast) {
    __syncthreads();
    build_lists_exact(d, P, s, tid, pi, cand, cflag, nc, stage, !lj_fast, !q_fast, cnt_lj, cnt_q);
  }
  if (active) {
    const int qi = d.topo.qslot[tid];
    d.lj_cnt[gi] = cnt_lj;
    if (qi >= 0) d.q_cnt[s * nq + qi] = cnt_q;
    d.xb[gi] = make_float4((float)pi.x, (float)pi.y, (float)pi.z, 0.0f);
    if (cnt_lj > d.lj_cap) atomicMax(&S.need_lj, cnt_lj);
    if (cnt_q > d.q_cap) atomicMax(&S.need_q, cnt_q);
    const double dx = min_img(pi.x - cs.x, box, hbox), dy = min_img(pi.y - cs.y, box, hbox),
                 dz = min_img(pi.z - cs.z, box, hbox);
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > P.r_bound2 || (qi >= 0 && d2 > P.r_bound_q2)) S.deformed = 1;
  }
  // list sizes of this subunit for the statistics
  const double wl = wave_total(active ? (double)cnt_lj : 0.0), wq = wave_total(active ? (double)cnt_q : 0.0);
  __syncthreads();
  double* red = reinterpret_cast<double*>(stage);
  if ((tid & 63) == 0) { red[2 * (tid >> 6)] = wl; red[2 * (tid >> 6) + 1] = wq; }
  __syncthreads();
  if (tid == 0) {
    d.part[7 * (size_t)d.Ns + s] = red[0] + red[2];
    d.part[8 * (size_t)d.Ns + s] = red[1] + red[3];
  }
}

// ------------------------------------------------------------------------------------------
// Cell-list build (cut Yukawa). Each replica's box is cut into g^3 cells per grid (g = 1 or
// g >= 3; cell edge >= the list radius): one grid for all beads (LJ/WCA lists) and one for the
// charged beads (Yukawa lists). A fill kernel drops every bead of a rebuilding replica into its
// cell as an FP32 record; a query kernel, one workgroup per subunit, scans the 27 cells around
// each bead with the FP32 list radius plus 1e-3 sigma (lists are supersets of the exact ones;
// entries beyond a cutoff contribute nothing) and keeps each list sorted by bead id, the order
// in which the reference's update_pairlist appends (src/functions.cpp:93-110), so that forces
// do not depend on when lists were built.

__device__ __forceinline__ float wrapf(float x, float box, float hbox) {
  return x - box * floorf((x + hbox) / box);
}

__device__ __forceinline__ int cell_of(float x, float hbox, float inv_a, int g) {
  const int c = (int)((x + hbox) * inv_a);
  return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

__global__ void k_cells_clear(Dev d) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < d.ncell_lj; c += gridDim.x * blockDim.x) d.cnt_lj[c] = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < d.ncell_q; c += gridDim.x * blockDim.x) d.cnt_q[c] = 0;
}

constexpr int kFillSlots = 256;  // LDS hash of the distinct cells of one subunit's beads (> kBlock)

// Records of the subunit's beads into one grid: beads that share a cell are counted in an LDS
// hash (atomicCAS on the key, linear probing; the table is never full, so every probe ends),
// then one global atomic per distinct cell reserves its slots and each bead writes at its rank.
__device__ __forceinline__ void fill_grid(bool in, int c, const float4& rec, int* cnt, float4* cells, int cap, int* need,
                                          int* key, int* cntl, int* base, int tid) {
  for (int i = tid; i < kFillSlots; i += kBlock) { key[i] = -1; cntl[i] = 0; }
  __syncthreads();
  int h = 0, rank = 0;
  if (in) {
    h = (int)(((unsigned)c * 2654435761u) >> 24);
    while (true) {
      const int k = atomicCAS(&key[h], -1, c);
      if (k == -1 || k == c) break;
      h = (h + 1) & (kFillSlots - 1);
    }
    rank = atomicAdd(&cntl[h], 1);
  }
  __syncthreads();
  for (int i = tid; i < kFillSlots; i += kBlock)
    if (key[i] >= 0) {
      base[i] = atomicAdd(&cnt[key[i]], cntl[i]);
      if (base[i] + cntl[i] > cap) atomicMax(need, base[i] + cntl[i]);
    }
  __syncthreads();
  if (in) {
    const int slot = base[h] + rank;
    if (slot < cap) st_stream(&cells[(size_t)c * cap + slot], rec);
  }
  __syncthreads();  // the tables are reused by the next grid
}

// One workgroup per subunit of a rebuilding replica: its beads into the LJ grid and its charged
// beads into the Yukawa grid, as FP32 records (x, y, z, id) of the wrapped positions.
__global__ __launch_bounds__(kBlock) void k_cell_fill(Dev d, int force_rebuild) {
  __shared__ int key[kFillSlots], cntl[kFillSlots], base[kFillSlots];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = rep_of(d, s);
  RepState& S = d.rs[r];
  if (!force_rebuild && S.rebuild == 0) return;
  const RepParams& P = d.rp[r];
  const int np = d.topo.np;
  const bool active = tid < np;
  const int b = s * np + tid;
  const float box = (float)P.box, hbox = (float)P.hbox;
  float4 rec = make_float4(0, 0, 0, 0);
  int clj = 0, cq = 0;
  bool charged = false;
  if (active) {
    const double4 p = d.pos[b];
    const float x = wrapf((float)p.x, box, hbox), y = wrapf((float)p.y, box, hbox), z = wrapf((float)p.z, box, hbox);
    rec = make_float4(x, y, z, __int_as_float(b));
    int g = P.g_lj;
    clj = P.c0_lj + (cell_of(z, hbox, P.inv_a_lj, g) * g + cell_of(y, hbox, P.inv_a_lj, g)) * g +
          cell_of(x, hbox, P.inv_a_lj, g);
    charged = d.topo.qslot[tid] >= 0;
    g = P.g_q;
    cq = P.c0_q + (cell_of(z, hbox, P.inv_a_q, g) * g + cell_of(y, hbox, P.inv_a_q, g)) * g + cell_of(x, hbox, P.inv_a_q, g);
  }
  fill_grid(active, clj, rec, d.cnt_lj, d.cell_lj, d.ccap_lj, &S.need_cell_lj, key, cntl, base, tid);
  if (d.topo.nq > 0)
    fill_grid(active && charged, cq, rec, d.cnt_q, d.cell_q, d.ccap_q, &S.need_cell_q, key, cntl, base, tid);
}

constexpr int kCqRecords = 512;   // LDS records (other subunits) of the cells around one subunit
constexpr int kCqCells = 512;     // cells of the box around one subunit's stencils
constexpr int kCqLjSlots = 12;    // LDS slots per bead for its sorted LJ list
constexpr int kCqQMax = 8;        // charged beads per template with LDS Yukawa slots
constexpr int kCqQSlots = 128;    // LDS slots per charged bead for its Yukawa list (a multiple of 4)
static_assert(kCqQMax * kCqQSlots * 4 <= kCqRecords * 16, "Yukawa slots live in the record space");
static_assert(3 * kCqCells <= kCqLjSlots * kBlock, "the gather's per-cell tables live in the LJ slots");
static_assert(kCqCells == 4 * kBlock, "the gather gives each thread four cells");

// In-place insertion sort of one list column (stride n) in global memory; used only when a
// list outgrew its LDS slots.
__device__ void sort_column(int* a, size_t stride, int n) {
  for (int k = 1; k < n; k++) {
    const int v = a[(size_t)k * stride];
    int p = k;
    while (p > 0 && a[(size_t)(p - 1) * stride] > v) {
      a[(size_t)p * stride] = a[(size_t)(p - 1) * stride];
      p--;
    }
    a[(size_t)p * stride] = v;
  }
}

__device__ __forceinline__ int wrap_cell(int c, int g) { return c < 0 ? c + g : (c >= g ? c - g : c); }

// Signed cell offset of c from cc on a periodic axis of g cells (nearest image).
__device__ __forceinline__ int rel_cell(int c, int cc, int g) {
  int r = c - cc;
  if (2 * r > g) r -= g;
  else if (2 * r < -g) r += g;
  return r;
}

// LJ / WCA list radius^2 test of bead i against candidate record e (FP32, +1e-3 sigma).
struct LjTest {
  float ix, iy, iz, si, ratt2, skinf, box, hbox, ibox;
  unsigned long long mask_i;
  __device__ __forceinline__ float r2_to(const float4& e) const {
    const float dx = min_imgf_near(ix - e.x, box, ibox), dy = min_imgf_near(iy - e.y, box, ibox),
                dz = min_imgf_near(iz - e.z, box, ibox);
    return dx * dx + dy * dy + dz * dz;
  }
  // the same for a record known to lie less than half a box away on every axis, where the shift
  // of min_imgf_near rounds to zero: the same value
  template <bool WRAP>
  __device__ __forceinline__ float r2(const float4& e) const {
    if (WRAP) return r2_to(e);
    const float dx = ix - e.x, dy = iy - e.y, dz = iz - e.z;
    return dx * dx + dy * dy + dz * dz;
  }
  __device__ __forceinline__ bool hit(float r2, int tj, float sj) const {
    if ((mask_i >> tj) & 1ull) return r2 < ratt2;
    const float w = (float)kWcaFactor * 0.5f * (si + sj) + skinf;
    return r2 < w * w;
  }
  // the largest list radius^2 of this bead over partners of sigma <= smax (a bound for hit())
  __device__ __forceinline__ float bound2(float smax) const {
    const float w = (float)kWcaFactor * 0.5f * (si + smax) + skinf;
    return mask_i ? fmaxf(ratt2, w * w) : w * w;
  }
  __device__ __forceinline__ bool operator()(const float4& e, int tj, float sj) const { return hit(r2_to(e), tj, sj); }
};

// Local box of cells around a subunit on one axis: cells cc + lo .. cc + lo + span - 1
// (periodic), or the whole axis when the stencils wrap around it (full).
struct CellAxis {
  int cc, lo, span, g;
  bool full;
  __device__ __forceinline__ int local(int rel) const { return full ? wrap_cell(cc + rel, g) : rel - lo; }
  __device__ __forceinline__ int global(int l) const { return full ? l : wrap_cell(cc + lo + l, g); }
  // periodic image of local cell l (or of relative cell rel) in the workgroup's unwrapped frame:
  // 1 = one box below the stored position, 2 = one box above, 0 = as stored (always for full axes)
  __device__ __forceinline__ int shift_code(int l) const {
    if (full) return 0;
    const int u = cc + lo + l;
    return u < 0 ? 1 : (u >= g ? 2 : 0);
  }
  __device__ __forceinline__ float shift_rel(int rel, float box) const {
    if (full) return 0.0f;
    const int u = cc + rel;
    return u < 0 ? -box : (u >= g ? box : 0.0f);
  }
};
__device__ __forceinline__ float shift_of(int code, float box) { return code == 1 ? -box : (code == 2 ? box : 0.0f); }

// The cells of a g^3 grid (first cell c0) that the 27-cell stencils of a workgroup's
// participating beads touch are marked, and the records of other subunits in them (the cell's
// records less the workgroup's own beads, ids in [excl_lo, excl_hi), whose cells the
// participating beads count themselves) are copied into LDS packed in local-cell order: cell lc
// holds rec[coff[lc] .. coff[lc] + cnum[lc]), and consecutive cells are contiguous. Records are
// moved to their periodic image in the workgroup's unwrapped frame around the centre cell (along
// axes the local box does not cover whole), so that beads moved into the same frame take plain
// differences. Each thread owns four consecutive local cells for the counting, the prefix sum and
// the copy. Returns false (uniformly) when the box or the records do not fit.
__device__ __forceinline__ bool gather_cells(const float4* cells, const int* cnt, int cap, int c0, int g, float inv_a, float hbox,
                             float cx, float cy, float cz, bool part, float px, float py, float pz, int tid,
                             CellAxis* ax, int* rel_out, int* cnum, int* ownc, int* coff, float4* rec, int* red_i,
                             int excl_lo, int excl_hi, float box, int reach) {
  const int ccx = cell_of(cx, hbox, inv_a, g), ccy = cell_of(cy, hbox, inv_a, g), ccz = cell_of(cz, hbox, inv_a, g);
  int rx = 0, ry = 0, rz = 0;
  if (part) {
    rx = rel_cell(cell_of(px, hbox, inv_a, g), ccx, g);
    ry = rel_cell(cell_of(py, hbox, inv_a, g), ccy, g);
    rz = rel_cell(cell_of(pz, hbox, inv_a, g), ccz, g);
  }
  rel_out[0] = rx; rel_out[1] = ry; rel_out[2] = rz;
  // block min / max of the relative cells (non-participants contribute the centre cell)
  int mn[3] = {rx, ry, rz}, mx[3] = {rx, ry, rz};
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = min(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = max(mx[k], __shfl_xor(mx[k], o, 64));
    }
  if ((tid & 63) == 0)
    for (int k = 0; k < 3; k++) { red_i[6 * (tid >> 6) + k] = mn[k]; red_i[6 * (tid >> 6) + 3 + k] = mx[k]; }
  __syncthreads();
  if (tid < 3) {
    const int lo = min(red_i[tid], red_i[6 + tid]) - 1, hi = max(red_i[3 + tid], red_i[9 + tid]) + 1;
    CellAxis& A = ax[tid];
    A.g = g;
    A.cc = tid == 0 ? ccx : (tid == 1 ? ccy : ccz);
    if (g == 1) { A.full = true; A.lo = 0; A.span = 1; }
    else if (hi - lo + 1 >= g) { A.full = true; A.lo = 0; A.span = g; }
    else { A.full = false; A.lo = lo; A.span = hi - lo + 1; }
  }
  __syncthreads();
  const int sx = ax[0].span, sy = ax[1].span, sz = ax[2].span;
  const int total = sx * sy * sz;
  if (total > kCqCells) return false;  // uniform: every thread reads the same spans
  for (int lc = tid; lc < total; lc += kBlock) { cnum[lc] = 0; ownc[lc] = 0; }
  __syncthreads();
  if (part) {  // mark the stencil cells the bead reaches (cnum used as a flag first) and count the own bead
    const bool one = g == 1;
    const int z0 = (!one && (reach & 4)) ? -1 : 0, z1 = (!one && (reach & 32)) ? 1 : 0;
    const int y0 = (!one && (reach & 2)) ? -1 : 0, y1 = (!one && (reach & 16)) ? 1 : 0;
    const int x0 = (!one && (reach & 1)) ? -1 : 0, x1 = (!one && (reach & 8)) ? 1 : 0;
    for (int dz = z0; dz <= z1; dz++) {
      const int lz = ax[2].local(rz + dz);
      for (int dy = y0; dy <= y1; dy++) {
        const int ly = ax[1].local(ry + dy);
        for (int dx = x0; dx <= x1; dx++) cnum[(lz * sy + ly) * sx + ax[0].local(rx + dx)] = 1;
      }
    }
    atomicAdd(&ownc[(ax[2].local(rz) * sy + ax[1].local(ry)) * sx + ax[0].local(rx)], 1);
  }
  __syncthreads();
  // records of other subunits per marked cell, then exclusive prefix sums of those (the packed
  // LDS layout) and of all stored records (the copy's work list)
  int loc[4], src[4], gcell[4], sum = 0, ssum = 0;
  // local coordinates of the thread's first cell; the next three follow by counting
  int lz = (4 * tid) / (sx * sy), ly = (4 * tid - lz * sx * sy) / sx, lx = 4 * tid - (lz * sy + ly) * sx;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int lc = 4 * tid + k;
    int n = 0, sn = 0, gl = c0;
    int code = 0;
    if (k > 0 && ++lx == sx) {
      lx = 0;
      if (++ly == sy) { ly = 0; lz++; }
    }
    if (lc < total && cnum[lc]) {
      gl = c0 + (ax[2].global(lz) * g + ax[1].global(ly)) * g + ax[0].global(lx);
      code = ax[0].shift_code(lx) | (ax[1].shift_code(ly) << 2) | (ax[2].shift_code(lz) << 4);
      sn = min(cnt[gl], cap);
      n = max(sn - ownc[lc], 0);
    }
    // global cell (< 2^24: grids hold at most two cells per bead, beads < 2^23) and image code
    loc[k] = n; src[k] = sn; gcell[k] = gl | (code << 24);
    sum += n; ssum += sn;
  }
  int incl = sum, sincl = ssum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64), sv = __shfl_up(sincl, o, 64);
    if ((tid & 63) >= o) { incl += v; sincl += sv; }
  }
  __syncthreads();
  if ((tid & 63) == 63) { red_i[tid >> 6] = incl; red_i[2 + (tid >> 6)] = sincl; }
  __syncthreads();
  const int base = incl - sum + ((tid >> 6) ? red_i[0] : 0);
  const int sbase = sincl - ssum + ((tid >> 6) ? red_i[2] : 0);
  const int all = red_i[0] + red_i[1], all_src = red_i[2] + red_i[3];
  // the own-bead counts are no longer needed: their space holds, per local cell, the end of its
  // stored records in the work list, its global cell and a running count of kept records
  int* send = ownc;
  int* gcl = ownc + kCqCells;
  int* run = ownc + 2 * kCqCells;
  int acc = base, sacc = sbase;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int lc = 4 * tid + k;
    coff[lc] = acc;
    cnum[lc] = loc[k];
    acc += loc[k];
    sacc += src[k];
    send[lc] = sacc;
    gcl[lc] = gcell[k];
    run[lc] = 0;
  }
  if (tid == 0) coff[kCqCells] = all;
  if (all > kCqRecords) return false;  // uniform
  __syncthreads();
  // copy, one stored record per thread and round: its cell is the first with send > rr
  for (int rr = tid; rr < all_src; rr += kBlock) {
    int lo = 0, hi = kCqCells - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (send[mid] > rr) hi = mid; else lo = mid + 1;
    }
    const int start = lo ? send[lo - 1] : 0;
    const int gc = gcl[lo];
    float4 v = cells[(size_t)(gc & 0xffffff) * cap + (rr - start)];
    const int j = __float_as_int(v.w);
    if (j >= excl_lo && j < excl_hi) continue;
    const int code = gc >> 24;  // < 64: the packed value stays positive
    v.x += shift_of(code & 3, box);
    v.y += shift_of((code >> 2) & 3, box);
    v.z += shift_of(code >> 4, box);
    const int w = atomicAdd(&run[lo], 1);
    if (w < cnum[lo]) rec[coff[lo] + w] = v;
  }
  __syncthreads();
  return true;
}

// The 27 cells around (cx, cy, cz) of a g^3 grid starting at c0 (periodic; g >= 3), or the
// single cell (g = 1).
template <typename F>
__device__ __forceinline__ void for_stencil(int c0, int g, int cx, int cy, int cz, F&& f) {
  if (g == 1) { f(c0); return; }
  for (int dz = -1; dz <= 1; dz++) {
    const int zz = wrap_cell(cz + dz, g);
    for (int dy = -1; dy <= 1; dy++) {
      const int yy = wrap_cell(cy + dy, g);
      for (int dx = -1; dx <= 1; dx++) f(c0 + (zz * g + yy) * g + wrap_cell(cx + dx, g));
    }
  }
}

// Lists of bead gi (active thread) straight from the global cells, written unsorted and then
// sorted in place: the path of a workgroup whose neighbourhood does not fit in LDS.
__device__ __forceinline__ void cell_query_direct(const Dev& d, const RepParams& P, int s, int gi, int tid, const LjTest T,
                                  const float2* tts, int* counts) {
  int cnt_lj = 0, cnt_q = 0;
  const int np = d.topo.np, nq = d.topo.nq;
  const float inv_np = 1.0f / (float)np;
  cnt_lj = 0;
  {
    const int g = P.g_lj;
    for_stencil(P.c0_lj, g, cell_of(T.ix, T.hbox, P.inv_a_lj, g), cell_of(T.iy, T.hbox, P.inv_a_lj, g),
                cell_of(T.iz, T.hbox, P.inv_a_lj, g), [&](int c) {
      const int n = min(d.cnt_lj[c], d.ccap_lj);
      const float4* cell = d.cell_lj + (size_t)c * d.ccap_lj;
      for (int m = 0; m < n; m++) {
        const float4 e = cell[m];
        const int j = __float_as_int(e.w);
        const int lj = local_of_small(j, np, inv_np);
        if (j - lj == s * np || !T(e, __float_as_int(tts[lj].x), tts[lj].y)) continue;
        if (cnt_lj < d.lj_cap) d.lj_list[(size_t)cnt_lj * d.Nb + gi] = j;
        cnt_lj++;
      }
    });
    sort_column(d.lj_list + gi, d.Nb, min(cnt_lj, d.lj_cap));
  }
  cnt_q = 0;
  const int qi = d.topo.qslot[tid];
  if (qi >= 0) {
    const int gq = s * nq + qi;
    const float rq = sqrtf((float)P.rc_q_list2) + 1e-3f;
    const float rq2 = rq * rq;
    const int g = P.g_q;
    for_stencil(P.c0_q, g, cell_of(T.ix, T.hbox, P.inv_a_q, g), cell_of(T.iy, T.hbox, P.inv_a_q, g),
                cell_of(T.iz, T.hbox, P.inv_a_q, g), [&](int c) {
      const int n = min(d.cnt_q[c], d.ccap_q);
      const float4* cell = d.cell_q + (size_t)c * d.ccap_q;
      for (int m = 0; m < n; m++) {
        const float4 e = cell[m];
        const int j = __float_as_int(e.w);
        const float dx = min_imgf_near(T.ix - e.x, T.box, T.ibox), dy = min_imgf_near(T.iy - e.y, T.box, T.ibox),
                    dz = min_imgf_near(T.iz - e.z, T.box, T.ibox);
        if (j == gi || !(dx * dx + dy * dy + dz * dz < rq2)) continue;
        if (cnt_q < d.q_cap) d.q_list[(size_t)cnt_q * d.Nq + gq] = j;
        cnt_q++;
      }
    });
    sort_column(d.q_list + gq, d.Nq, min(cnt_q, d.q_cap));
  }
  counts[0] = cnt_lj;
  counts[1] = cnt_q;
}

// One workgroup per subunit of a rebuilding replica: the cells its beads' stencils touch are
// gathered into LDS (all beads for the LJ lists, then the charged beads for the Yukawa lists);
// each bead keeps its LJ hits insertion-sorted in LDS, the Yukawa hits are ranked by all
// threads together. Neighbourhoods that do not fit in LDS take the direct path.
__global__ __launch_bounds__(kBlock) void k_cell_query(Dev d, int force_rebuild) {
  // LDS shared between phases: the Yukawa slots reuse the records once the LJ lists are out,
  // and the own-bead cell counts of the gather the LJ slots before the LJ scan
  __shared__ float4 rec[kCqRecords];
  __shared__ int cnum[kCqCells];
  __shared__ int coff[kCqCells + 1];
  __shared__ CellAxis ax[3];
  __shared__ int red_i[16];
  __shared__ int ljbuf[kCqLjSlots * kBlock];
  int* const ownc = ljbuf;
  int* const qbuf = reinterpret_cast<int*>(rec);
  __shared__ int qn[kCqQMax];
  __shared__ int qoff[kCqQMax + 1];
  __shared__ float4 qsh[kCqQMax];
  __shared__ int qcell[kCqQMax][3];
  __shared__ float2 tts[kBlock];  // (type bits, sigma) of each template bead, one load per record
  __shared__ float smax_w[2];
  __shared__ unsigned long long tmask[kMaxTypes];
  __shared__ double red[4];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = rep_of(d, s);
  RepState& S = d.rs[r];
  if (!force_rebuild && S.rebuild == 0) return;
  const RepParams& P = d.rp[r];
  const int np = d.topo.np, nq = d.topo.nq;
  const bool active = tid < np;
  const int gi = s * np + tid;
  if (active) tts[tid] = make_float2(__int_as_float(d.topo.type[tid]), (float)d.topo.sigma[tid]);
  if (tid < kMaxTypes) tmask[tid] = d.topo.att_mask[tid];
  if (tid < kCqQMax) qn[tid] = 0;
  const float box = (float)P.box, hbox = (float)P.hbox;
  const float inv_np = 1.0f / (float)np;
  double4 pi = make_double4(0, 0, 0, 0);
  if (active) pi = d.pos[gi];
  const double4 cs = d.centres[s];
  const float cx = wrapf((float)cs.x, box, hbox), cy = wrapf((float)cs.y, box, hbox), cz = wrapf((float)cs.z, box, hbox);
  LjTest T;
  T.ix = wrapf((float)pi.x, box, hbox); T.iy = wrapf((float)pi.y, box, hbox); T.iz = wrapf((float)pi.z, box, hbox);
  T.box = box; T.hbox = hbox; T.ibox = (float)P.ibox;
  T.si = active ? (float)d.topo.sigma[tid] : 0.0f;
  T.mask_i = active ? d.topo.att_mask[d.topo.type[tid]] : 0ull;
  const float ratt = (float)(P.ecut + P.skin) + 1e-3f;
  T.ratt2 = ratt * ratt;
  T.skinf = (float)P.skin + 1e-3f;
  {  // largest template sigma, for the cheap distance filter of the LJ scan
    float sm = T.si;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sm = fmaxf(sm, __shfl_xor(sm, o, 64));
    if ((tid & 63) == 0) smax_w[tid >> 6] = sm;
  }
  __syncthreads();
  const float bound2 = T.bound2(fmaxf(smax_w[0], smax_w[1]));
  int cnt_lj = 0, cnt_q = 0;
  int rel[3];
  // ---- LJ / WCA: the neighbourhood's records of other subunits, gathered into LDS.
  // The neighbour cells this bead's list radius reaches: on each axis the lower (upper) one only
  // when the radius box crosses the lower (upper) face of the bead's cell (bits k and 3 + k);
  // records in the other cells fail the radius test anyway (a WCA-only bead, radius ~1.4 sigma
  // against cells of the largest list radius, needs at most two cells per axis). Only reached
  // cells are gathered and scanned.
  int reach = 0;
  {
    const float rr = sqrtf(bound2) + 1e-3f, ia = P.inv_a_lj;
    const float pc[3] = {T.ix, T.iy, T.iz};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float tc = (pc[k] + hbox) * ia, fl = floorf(tc), f = tc - fl;
      const bool in = fl >= 0.0f && fl < (float)P.g_lj;  // else the cell was clamped: both neighbours
      if (!in || f < rr * ia) reach |= 1 << k;
      if (!in || 1.0f - f < rr * ia) reach |= 8 << k;
    }
  }
  const bool lj_ok = gather_cells(d.cell_lj, d.cnt_lj, d.ccap_lj, P.c0_lj, P.g_lj, P.inv_a_lj, hbox, cx, cy, cz, active,
                                  T.ix, T.iy, T.iz, tid, ax, rel, cnum, ownc, coff, rec, red_i, s * np, (s + 1) * np, box,
                                  reach);
  bool direct = !lj_ok;
  if (lj_ok && active) {
    // the local box in registers; with no wrap along x the three cells of a (dz, dy) row are one
    // contiguous record range, otherwise each of the 27 cells is its own range
    const int g = ax[0].g, sx = ax[0].span, sy = ax[1].span;
    const bool fx = ax[0].full, fy = ax[1].full, fz = ax[2].full;
    const int ox = fx ? ax[0].cc : -ax[0].lo, oy = fy ? ax[1].cc : -ax[1].lo, oz = fz ? ax[2].cc : -ax[2].lo;
    const bool rows = !fx && g > 1;
    const int nseg = g == 1 ? 1 : (rows ? 9 : 27);
    bool lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = (reach >> k) & 1;
      hi[k] = (reach >> (3 + k)) & 1;
    }
    // the bead in the records' unwrapped frame; no record then needs an image shift unless the
    // local box covers an axis whole or spans half of it
    LjTest TL = T;
    TL.ix += ax[0].shift_rel(rel[0], box);
    TL.iy += ax[1].shift_rel(rel[1], box);
    TL.iz += ax[2].shift_rel(rel[2], box);
    bool wrap = g == 1;
#pragma unroll
    for (int k = 0; k < 3; k++) wrap = wrap || ax[k].full || 2 * ax[k].span >= g;
    bool over = false;
    auto scan = [&](auto wrap_tag) {
    constexpr bool W = decltype(wrap_tag)::value;
    for (int pass = 0; pass < 2; pass++) {  // pass 1 only when the LDS slots overflowed
      if (pass == 1) {
        if (!over) break;
        cnt_lj = 0;
      }
      for (int sg = 0; sg < nseg; sg++) {
        int m0, m1;
        if (g == 1) {
          m0 = coff[0]; m1 = m0 + cnum[0];
        } else {
          const int dz = rows ? sg / 3 - 1 : sg / 9 - 1, dy = rows ? sg % 3 - 1 : (sg / 3) % 3 - 1;
          if ((dz < 0 && !lo[2]) || (dz > 0 && !hi[2]) || (dy < 0 && !lo[1]) || (dy > 0 && !hi[1])) continue;
          const int lz = fz ? wrap_cell(oz + rel[2] + dz, g) : rel[2] + dz + oz;
          const int ly = fy ? wrap_cell(oy + rel[1] + dy, g) : rel[1] + dy + oy;
          if (rows) {
            const int l0 = (lz * sy + ly) * sx + rel[0] - 1 + ox;
            m0 = coff[l0 + (lo[0] ? 0 : 1)]; m1 = coff[l0 + (hi[0] ? 3 : 2)];
          } else {
            const int dx = sg % 3 - 1;
            if ((dx < 0 && !lo[0]) || (dx > 0 && !hi[0])) continue;
            const int lc = (lz * sy + ly) * sx + wrap_cell(ox + rel[0] + dx, g);
            m0 = coff[lc]; m1 = m0 + cnum[lc];
          }
        }
        for (int m = m0; m < m1; m += 4) {  // four records requested together
          float4 e[4];
          float r2[4];
#pragma unroll
          for (int u = 0; u < 4; u++) e[u] = rec[min(m + u, m1 - 1)];
          // distances first, without branches; the pair's type and sigma only for the few records
          // inside the bound, in record order
          unsigned cand = 0;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            r2[u] = TL.r2<W>(e[u]);
            cand |= (m + u < m1 && r2[u] < bound2) ? 1u << u : 0u;
          }
          if (cand == 0) continue;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (!((cand >> u) & 1u)) continue;
            const int j = __float_as_int(e[u].w);
            const float2 tj = tts[local_of_small(j, np, inv_np)];
            if (!TL.hit(r2[u], __float_as_int(tj.x), tj.y)) continue;
            if (pass == 1) {
              if (cnt_lj < d.lj_cap) d.lj_list[(size_t)cnt_lj * d.Nb + gi] = j;
            } else if (cnt_lj < kCqLjSlots) {
              int p = cnt_lj;
              while (p > 0 && ljbuf[(p - 1) * kBlock + tid] > j) {
                ljbuf[p * kBlock + tid] = ljbuf[(p - 1) * kBlock + tid];
                p--;
              }
              ljbuf[p * kBlock + tid] = j;
            } else {
              over = true;
            }
            cnt_lj++;
          }
        }
      }
    }
    };
    if (wrap) scan(Flag<true>{}); else scan(Flag<false>{});
    if (!over) {
      for (int k = 0; k < cnt_lj && k < d.lj_cap; k++) st_stream(&d.lj_list[(size_t)k * d.Nb + gi], ljbuf[k * kBlock + tid]);
    } else {  // rare: the list went to global memory in pass 1; sort it there
      sort_column(d.lj_list + gi, d.Nb, min(cnt_lj, d.lj_cap));
    }
  }
  // ---- Yukawa: charged beads j != i within ecut_el + skin. Work items (charged bead, stencil
  // cell) spread over all threads; hits are appended in LDS and ranked afterwards, so the
  // list order (ascending id) does not depend on the append order.
  if (nq > 0 && !direct) {
    __syncthreads();
    const int qi = active ? d.topo.qslot[tid] : -1;
    const int g = P.g_q;
    if (qi >= 0 && qi < kCqQMax) {
      qsh[qi] = make_float4(T.ix, T.iy, T.iz, __int_as_float(gi));
      qcell[qi][0] = cell_of(T.ix, hbox, P.inv_a_q, g);
      qcell[qi][1] = cell_of(T.iy, hbox, P.inv_a_q, g);
      qcell[qi][2] = cell_of(T.iz, hbox, P.inv_a_q, g);
    }
    __syncthreads();
    const float rq = sqrtf((float)P.rc_q_list2) + 1e-3f;
    const float rq2 = rq * rq;
    const int nql = min(nq, kCqQMax);
    const int ncell = g == 1 ? 1 : 27;
    // with five or more cells per axis the stencil's records all lie less than half a box from the
    // bead along each axis once the stencil cell's periodic image is applied to the bead (one box
    // when the cell index wrapped): plain differences then; smaller grids take the minimum image
    auto yscan = [&](auto wrap_tag) {
      constexpr bool W = decltype(wrap_tag)::value;
      for (int it = tid; it < nql * ncell; it += kBlock) {
        const int a = it / ncell, k = it - a * ncell;
        int c = P.c0_q;
        float4 qa = qsh[a];
        if (g > 1) {
          const int ux = qcell[a][0] + k % 3 - 1, uy = qcell[a][1] + (k / 3) % 3 - 1, uz = qcell[a][2] + k / 9 - 1;
          c += (wrap_cell(uz, g) * g + wrap_cell(uy, g)) * g + wrap_cell(ux, g);
          if (!W) {  // the bead's image next to the cell's records
            qa.x += ux < 0 ? box : (ux >= g ? -box : 0.0f);
            qa.y += uy < 0 ? box : (uy >= g ? -box : 0.0f);
            qa.z += uz < 0 ? box : (uz >= g ? -box : 0.0f);
          }
        }
        const int self = __float_as_int(qa.w);
        const int n = min(d.cnt_q[c], d.ccap_q);
        const float4* cell = d.cell_q + (size_t)c * d.ccap_q;
        for (int m0 = 0; m0 < n; m0 += 8) {  // eight records requested together
          float4 e[8];
#pragma unroll
          for (int u = 0; u < 8; u++)
            if (m0 + u < n) e[u] = cell[m0 + u];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            if (m0 + u >= n) break;
            const int j = __float_as_int(e[u].w);
            float dx = qa.x - e[u].x, dy = qa.y - e[u].y, dz = qa.z - e[u].z;
            if (W) {
              dx = min_imgf_near(dx, box, T.ibox);
              dy = min_imgf_near(dy, box, T.ibox);
              dz = min_imgf_near(dz, box, T.ibox);
            }
            if (j == self || !(dx * dx + dy * dy + dz * dz < rq2)) continue;
            const int slot = atomicAdd(&qn[a], 1);
            if (slot < kCqQSlots) qbuf[a * kCqQSlots + slot] = j;
          }
        }
      }
    };
    if (g >= 5) yscan(Flag<false>{}); else yscan(Flag<true>{});
    __syncthreads();
    // rank sort: an entry's rank is the number of smaller ids in its list; the entries of all
    // lists are spread over the threads (lists that overflowed their slots are left to the
    // direct path below)
    if (tid == 0) {
      int o = 0;
      for (int a = 0; a < nql; a++) {
        qoff[a] = o;
        o += qn[a] <= kCqQSlots ? qn[a] : 0;
      }
      qoff[nql] = o;
    }
    __syncthreads();
    for (int e = tid; e < qoff[nql]; e += kBlock) {
      int a = 0;
      while (a + 1 < nql && qoff[a + 1] <= e) a++;
      const int n = qn[a];
      const int* buf = qbuf + a * kCqQSlots;
      const int v = buf[e - qoff[a]];
      const int4* b4 = reinterpret_cast<const int4*>(buf);
      int rank = 0;
#pragma unroll 4
      for (int m = 0; m < (n >> 2); m++) {
        const int4 w = b4[m];
        rank += (w.x < v) + (w.y < v) + (w.z < v) + (w.w < v);
      }
      for (int m = n & ~3; m < n; m++) rank += buf[m] < v ? 1 : 0;
      if (rank < d.q_cap) st_stream(&d.q_list[(size_t)rank * d.Nq + s * nq + a], v);
    }
    if (qi >= 0) {
      const int gq = s * nq + qi;
      if (qi < kCqQMax && qn[qi] <= kCqQSlots) {
        cnt_q = qn[qi];
      } else {  // rare: this bead's list straight from the global cells, sorted in place
        cnt_q = 0;
        for_stencil(P.c0_q, g, cell_of(T.ix, hbox, P.inv_a_q, g), cell_of(T.iy, hbox, P.inv_a_q, g),
                    cell_of(T.iz, hbox, P.inv_a_q, g), [&](int c) {
          const int n = min(d.cnt_q[c], d.ccap_q);
          const float4* cell = d.cell_q + (size_t)c * d.ccap_q;
          for (int m = 0; m < n; m++) {
            const float4 e = cell[m];
            const int j = __float_as_int(e.w);
            const float dx = min_imgf_near(T.ix - e.x, box, T.ibox), dy = min_imgf_near(T.iy - e.y, box, T.ibox),
                        dz = min_imgf_near(T.iz - e.z, box, T.ibox);
            if (j == gi || !(dx * dx + dy * dy + dz * dz < rq2)) continue;
            if (cnt_q < d.q_cap) d.q_list[(size_t)cnt_q * d.Nq + gq] = j;
            cnt_q++;
          }
        });
        sort_column(d.q_list + gq, d.Nq, min(cnt_q, d.q_cap));
      }
      d.q_cnt[gq] = cnt_q;
      if (cnt_q > d.q_cap) atomicMax(&S.need_q, cnt_q);
    }
  }
  if (direct) {  // uniform: the neighbourhood did not fit in LDS
    __syncthreads();
    if (active) {
      int* cts = ljbuf + 2 * tid;  // two LDS words per thread, free by now
      cell_query_direct(d, P, s, gi, tid, T, tts, cts);
      cnt_lj = cts[0];
      cnt_q = cts[1];
      const int qi = d.topo.qslot[tid];
      if (qi >= 0) {
        d.q_cnt[s * nq + qi] = cnt_q;
        if (cnt_q > d.q_cap) atomicMax(&S.need_q, cnt_q);
      }
    }
  }
  if (active) {
    d.lj_cnt[gi] = cnt_lj;
    d.xb[gi] = make_float4((float)pi.x, (float)pi.y, (float)pi.z, 0.0f);
    if (cnt_lj > d.lj_cap) atomicMax(&S.need_lj, cnt_lj);
  }
  // list sizes of this subunit for the statistics
  const double wl = wave_total(active ? (double)cnt_lj : 0.0), wq = wave_total(active ? (double)cnt_q : 0.0);
  if ((tid & 63) == 0) { red[2 * (tid >> 6)] = wl; red[2 * (tid >> 6) + 1] = wq; }
  __syncthreads();
  if (tid == 0) {
    d.part[7 * (size_t)d.Ns + s] = red[0] + red[2];
    d.part[8 * (size_t)d.Ns + s] = red[1] + red[3];
  }
}

// ------------------------------------------------------------------------------------------
// Bending of one hinge, the reference's force (src/edge.cpp:69-188): Kb / (4 a0 a1) (grad P -
// S (grad a0 / a0 + grad a1 / a1)), where S = 4 |A0.A1| (its face flip makes the product of the
// two area vectors non-negative) and P is the unflipped product its gradient formulas
// differentiate. With e = x1-x0, a = x2-x0, b = x3-x0, u = e x a, v = e x b (|u| = 2 a0,
// |v| = 2 a1) and c = u.v/(|u||v|), P = -u.v, so the force is -Kb (grad c + (c + |c|) grad
// ln(|u||v|)): for c <= 0 (faces folded less than 90 degrees) the gradient of the energy
// Kb (1 - |c|); past 90 degrees it is not a gradient, and it is what the reference computes.
// Per hinge it is the cross-product form below with dc/du replaced by
// g = (v/|v| + |c| u/|u|)/|u| and dc/dv by h = (u/|u| + |c| v/|v|)/|v|: force on x2 = -Kb g x e,
// x3: -Kb h x e, x1: -Kb (a x g + b x h), x0: minus their sum.
// Writes the 4 bead forces (12 doubles) and returns c (the energy is Kb (1 - |c|)).
template <bool WRAP>
__device__ __forceinline__ double hinge_forces(const double4* own, int4 hg, double box, double ibox, double Kb,
                                               double* out) {
  const double4 p0 = own[hg.x], p1 = own[hg.y], p2 = own[hg.z], p3 = own[hg.w];
  const double ex = near_diff<WRAP>(p1.x - p0.x, box, ibox), ey = near_diff<WRAP>(p1.y - p0.y, box, ibox),
               ez = near_diff<WRAP>(p1.z - p0.z, box, ibox);
  const double ax = near_diff<WRAP>(p2.x - p0.x, box, ibox), ay = near_diff<WRAP>(p2.y - p0.y, box, ibox),
               az = near_diff<WRAP>(p2.z - p0.z, box, ibox);
  const double bx = near_diff<WRAP>(p3.x - p0.x, box, ibox), by = near_diff<WRAP>(p3.y - p0.y, box, ibox),
               bz = near_diff<WRAP>(p3.z - p0.z, box, ibox);
  const double ux = ey * az - ez * ay, uy = ez * ax - ex * az, uz = ex * ay - ey * ax;
  const double vx = ey * bz - ez * by, vy = ez * bx - ex * bz, vz = ex * by - ey * bx;
  const double iu = rsqrt(ux * ux + uy * uy + uz * uz), iv = rsqrt(vx * vx + vy * vy + vz * vz);
  const double c = (ux * vx + uy * vy + uz * vz) * iu * iv;
  const double cn = -fabs(c);  // = c for the usual hinge (c <= 0)
  const double gx = (vx * iv - ux * (cn * iu)) * iu, gy = (vy * iv - uy * (cn * iu)) * iu,
               gz = (vz * iv - uz * (cn * iu)) * iu;
  const double hx = (ux * iu - vx * (cn * iv)) * iv, hy = (uy * iu - vy * (cn * iv)) * iv,
               hz = (uz * iu - vz * (cn * iv)) * iv;
  const double k = -Kb;
  const double f2x = (gy * ez - gz * ey) * k, f2y = (gz * ex - gx * ez) * k, f2z = (gx * ey - gy * ex) * k;
  const double f3x = (hy * ez - hz * ey) * k, f3y = (hz * ex - hx * ez) * k, f3z = (hx * ey - hy * ex) * k;
  const double f1x = ((ay * gz - az * gy) + (by * hz - bz * hy)) * k;
  const double f1y = ((az * gx - ax * gz) + (bz * hx - bx * hz)) * k;
  const double f1z = ((ax * gy - ay * gx) + (bx * hy - by * hx)) * k;
  out[0] = -(f1x + f2x + f3x); out[1] = -(f1y + f2y + f3y); out[2] = -(f1z + f2z + f3z);
  out[3] = f1x; out[4] = f1y; out[5] = f1z;
  out[6] = f2x; out[7] = f2y; out[8] = f2z;
  out[9] = f3x; out[10] = f3y; out[11] = f3z;
  return c;
}


// Per-template-bead tables a workgroup keeps in LDS for the neighbour lookups.
struct BeadLds {
  double sigma[kBlock];
  int type[kBlock];
  int qslot[kBlock];
  unsigned long long mask[kMaxTypes];
};

// LJ / WCA term on bead i from bead j (src/forces.cpp:112-139): attractive rows use
// 48 eps (s_i^6/r^6)((s_i^6/r^6) - 1/2)/r^2 below ecut with eps = elj_att, or 50 when
// type_j == 7; other pairs the WCA form with eps = 1 and s = (s_i + s_j)/2 below
// 1.12246205 s. Returns the scalar f of F = f r_vec (0 beyond the cutoff).
template <typename PT>
__device__ __forceinline__ PT lj_term(PT r2, bool att, int tj, PT si, PT sj, PT ecut2, PT elj) {
  const PT shc = (si + sj) / (PT)2;
  PT s6, eps;
  bool on;
  if (att) {
    on = r2 < ecut2;
    s6 = si * si * si * si * si * si;
    eps = (tj == 7) ? (PT)50 : elj;
  } else {
    on = r2 < (PT)(kWcaFactor * kWcaFactor) * shc * shc;
    s6 = shc * shc * shc * shc * shc * shc;
    eps = (PT)1;
  }
  if (!on) return (PT)0;
  const PT ir2 = (PT)1 / r2;  // one division; (s/r)^6 = s^6 (1/r^2)^3
  const PT sr6 = s6 * (ir2 * ir2 * ir2);
  return (PT)48 * eps * (sr6 * (sr6 - (PT)0.5)) * ir2;
}

// Screened Coulomb term (src/forces.cpp:100-108): q_i q_j lb exp(-k r)/r^2 (k + 1/r) Y_ij below
// ecut_el; returns f of F = f r_vec.
template <typename PT>
__device__ __forceinline__ PT yk_term(PT r2, PT qq, PT yk, PT kappa, PT ecut_el2) {
  if (!(r2 < ecut_el2)) return (PT)0;
  const PT ir = rsqrt(r2), rr = r2 * ir;  // no division: 1/r = ir, 1/r^2 = ir^2
  return (qq * exp(-kappa * rr) * (ir * ir)) * (kappa + ir) * yk;
}

// What one pair-list entry of a force workgroup needs besides its neighbour's position: the
// round's owner map and the per-bead offsets in LDS, the template tables, the subunit's own
// positions, the Yukawa tables and the cutoffs (passed by value; everything is uniform).
template <typename PT>
struct PairCtx {
  const unsigned char* owner;
  const int* off;
  const int* nlj;
  const BeadLds* tb;
  const double4* own;
  const double* qq;
  const double* yk;
  const double4* centres;
  double4 cs;
  double box, hbox, ibox, far_rf2;
  PT ecut2, elj, kappa, ecut_el2;
  int nq, split;
};

// Pair term of entry c0 + le of the subunit's concatenated lists (LJ entries, then Yukawa
// entries, bead by bead) on its bead from the neighbour at pj (src/forces.cpp:67-139); the
// force components go to pf[le], pf[kChunk + le], pf[2 kChunk + le].
template <typename PT, int kChunk, bool WRAP>
__device__ __forceinline__ void pair_entry(int le, int c0, double4 pj, PairCtx<PT> c, PT* pf) {
  const int i = c.owner[le];
  const int k = c0 + le - c.off[i];
  const int nl = c.nlj[i];
  const double4 p0 = c.own[i];
  const PT dx = (PT)near_diff<WRAP>(p0.x - pj.x, c.box, c.ibox), dy = (PT)near_diff<WRAP>(p0.y - pj.y, c.box, c.ibox),
           dz = (PT)near_diff<WRAP>(p0.z - pj.z, c.box, c.ibox);
  const PT r2 = dx * dx + dy * dy + dz * dz;
  const int lj = tag_loc(pj.w);
  PT f;
  if (k < nl) {
    const int tj = c.tb->type[lj];
    f = lj_term<PT>(r2, (c.tb->mask[c.tb->type[i]] >> tj) & 1ull, tj, (PT)c.tb->sigma[i], (PT)c.tb->sigma[lj], c.ecut2,
                    c.elj);
  } else {
    const int qs = c.tb->qslot[i], qj = c.tb->qslot[lj];
    f = yk_term<PT>(r2, (PT)c.qq[qs * c.nq + qj], (PT)c.yk[qs * c.nq + qj], c.kappa, c.ecut_el2);
    if (c.split) {  // uncut Yukawa: only subunit pairs within far_rf are near field
      const double4 cb = c.centres[tag_sub(pj.w)];
      const double cx = min_img(c.cs.x - cb.x, c.box, c.hbox), cy = min_img(c.cs.y - cb.y, c.box, c.hbox),
                   cz = min_img(c.cs.z - cb.z, c.box, c.hbox);
      if (cx * cx + cy * cy + cz * cz > c.far_rf2) f = (PT)0;
    }
  }
  pf[le] = dx * f;
  pf[kChunk + le] = dy * f;
  pf[2 * kChunk + le] = dz * f;
}

// LDS bytes of k_force: own positions, the hinge forces (three beads per hinge; the fourth is
// minus their sum) sharing their space with the pair terms of one gather round of `tile`
// entries (the bending sums are taken before the pair rounds), and the Yukawa tables.
static inline size_t force_lds_bytes(int np, int nh, int nq, int tile) {
  return (size_t)np * 32 + (size_t)std::max(nh * 9, tile * 3) * 8 + (size_t)nq * nq * 16;
}

// Encapsulation big beads on subunit bead p of template index l (src/forces.cpp:174-236 and
// src/energies.cpp:96-136), big beads in order. The immobile big beads exert a Yukawa force
// on charged beads within ecut_el and a WCA force shifted by delta = (sigma_big + sigma_j) / 2
// - shc; each term is subtracted from the bead's sums as there (the reference's vector is big
// bead minus subunit bead). ENERGY adds the energies (ne, ce) of the same terms.
template <bool ENERGY>
__device__ __forceinline__ void big_bead_terms(const Dev& d, const RepParams& P, double4 p, int l, double* lj,
                                               double* el, double* ne, double* ce) {
  const double box = P.box, hbox = P.hbox, kappa = P.kappa, ecut_el = P.ecut_el;
  const double4 t0 = d.big_tab[2 * l], t1 = d.big_tab[2 * l + 1];
  const double delta = t0.x, wca_cut = t0.y, wca_cut2 = t0.z, ys = t0.w;
  const double qqlb = t1.x, ecut_el2 = t1.z, shc6 = t1.w;
  const bool charged = t1.y != 0;
  for (int i = 0; i < d.nbig; i++) {
    const double4 c = d.big_pos[i];
    const double rx = min_img(c.x - p.x, box, hbox), ry = min_img(c.y - p.y, box, hbox),
                 rz = min_img(c.z - p.z, box, hbox);
    const double r2 = rx * rx + ry * ry + rz * rz;
    const bool yk = charged && r2 < ecut_el2, wca = r2 < wca_cut2;
    if (!(yk || wca || ENERGY)) continue;
    const double r = sqrt(r2);
    const double u = r - delta, u2 = u * u, r6 = u2 * u2 * u2;
    const double x6 = shc6 / r6;
    const double ek = charged && (yk || r < ecut_el) ? exp(-kappa * r) : 0.0;
    if (yk) {
      const double sc = ((qqlb * ek) / r2) * (kappa + 1 / r) * ys;
      el[0] -= rx * sc; el[1] -= ry * sc; el[2] -= rz * sc;
    }
    if (wca) {
      const double sc = 48 * (x6 * (x6 - 0.5)) * (1 / r) * (1 / u);
      lj[0] -= rx * sc; lj[1] -= ry * sc; lj[2] -= rz * sc;
    }
    if (ENERGY) {
      if (charged && r < ecut_el) *ce += (qqlb * ek * ys) / r;
      if (r < wca_cut) *ne += (4 * x6 * (x6 - 1)) + 1;
    }
  }
}

// kChunk: list entries the workgroup evaluates per gather round (the tile).
template <int MODE, typename PT, int kChunk>
// Four waves per SIMD is what the LDS allows (8 workgroups per CU): up to 128 VGPRs.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_force(Dev d, int init) {
  extern __shared__ double smem[];
  __shared__ BeadLds tb;
  __shared__ int off_s[kBlock + 1];
  __shared__ int nlj_s[kBlock];
  __shared__ unsigned char owner_s[kChunk];
  __shared__ int wflag[kBlock / 64];
  const int np = d.topo.np, nh = d.topo.nh, nq = d.topo.nq;
  double4* own = reinterpret_cast<double4*>(smem);
  double* hc = reinterpret_cast<double*>(own + np);          // [nh][9] forces on hinge beads 1-3
  PT* pf = reinterpret_cast<PT*>(hc);                        // [3][kChunk], after the bending sums
  double (*red)[kPartStride] = reinterpret_cast<double (*)[kPartStride]>(hc);  // epilogue sums (tile free)
  double* qq_s = hc + max(9 * nh, 3 * kChunk);
  double* yk_s = qq_s + nq * nq;

  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = rep_of(d, s);
  const RepParams& P = d.rp[r];
  RepState& S = d.rs[r];
  const double box = P.box, hbox = P.hbox, ibox = P.ibox, kb = P.kb, ks = P.ks, dt = P.dt;
  const double4 cs = d.centres[s];
  const bool active = tid < np;
  const int gi = s * np + tid;
  const int lt = active ? tid : 0;
  // one round of independent loads: own position, velocity, list sizes, the packed edge and
  // hinge rows, the template bead tables and the Yukawa tables
  int qi = -1, nlj = 0, nqe = 0;
  double4 p_own = cs;
  if (active) {
    p_own = d.pos[gi];
    own[tid] = p_own;
    tb.sigma[tid] = d.topo.sigma[tid];
    tb.type[tid] = d.topo.type[tid];
    qi = d.topo.qslot[tid];
    tb.qslot[tid] = qi;
    nlj = min(d.lj_cnt[gi], d.lj_cap);
    if (qi >= 0) nqe = min(d.q_cnt[s * nq + qi], d.q_cap);
  }
  if (tid < kMaxTypes) tb.mask[tid] = d.topo.att_mask[tid];
  for (int k = tid; k < nq * nq; k += kBlock) { qq_s[k] = P.qq[k]; yk_s[k] = P.yk[k]; }
  const int ecnt = d.topo.e_cnt[lt];
  const int hcnt = d.topo.h_cnt[lt];
  const int2 erow = reinterpret_cast<const int2*>(d.topo.e_other)[lt];
  const double2* lrow = reinterpret_cast<const double2*>(d.topo.e_len0) + (size_t)lt * (kMaxDeg / 2);
  double2 len[kMaxDeg / 2];
#pragma unroll
  for (int k = 0; k < kMaxDeg / 2; k++) len[k] = lrow[k];
  const int4 hrow0 = reinterpret_cast<const int4*>(d.topo.h_slot)[2 * lt];
  const int4 hrow1 = reinterpret_cast<const int4*>(d.topo.h_slot)[2 * lt + 1];
  const double m = d.topo.mass[lt];
  const int4 hg0 = reinterpret_cast<const int4*>(d.topo.hinge)[tid < nh ? tid : 0];
  const int4 hg1 = reinterpret_cast<const int4*>(d.topo.hinge)[tid + kBlock < nh ? tid + kBlock : 0];
  double vx = 0, vy = 0, vz = 0;
  if (active) { vx = d.vx[gi]; vy = d.vy[gi]; vz = d.vz[gi]; }
  const uint32_t skip = (d.bd_skip && active) ? d.bd_skip[gi - P.b0] : 1u;  // BD engine skip-ahead
  // offsets of each bead's entries in the subunit's concatenated lists (LJ entries, then
  // Yukawa entries, bead by bead): wave prefix sums, then the two waves in order
  const int n_i = nlj + nqe;
  int incl = n_i;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if ((tid & 63) >= o) incl += v;
  }
  if ((tid & 63) == 63) off_s[kBlock - 1 - (tid >> 6)] = incl;  // wave totals, parked at the end
  // differences that need no image shift (near_diff): the subunit's beads all within a quarter
  // box of its centre bead on every axis (bonded vectors), and, with positions kept inside the
  // box, all beads at least the list reach from the box faces (list neighbours, whose images are
  // then their raw positions)
  {
    const double q = 0.5 * hbox, lim = hbox - P.pair_reach;
    const bool ok_b = fabs(p_own.x - cs.x) < q && fabs(p_own.y - cs.y) < q && fabs(p_own.z - cs.z) < q;
    const bool ok_p = fabs(p_own.x) <= lim && fabs(p_own.y) <= lim && fabs(p_own.z) <= lim;
    const unsigned long long vb = __ballot(ok_b), vp = __ballot(ok_p);
    if ((tid & 63) == 0) wflag[tid >> 6] = (vb == ~0ull ? 1 : 0) | (vp == ~0ull ? 2 : 0);
  }
  __syncthreads();
  const int wf = __builtin_amdgcn_readfirstlane(wflag[0] & wflag[1]);
  const bool bond_wrap = !(wf & 1), pair_wrap = !(wf & 2) || !P.wrapped;
  const int wave0_total = off_s[kBlock - 1];
  const int off_i = incl - n_i + ((tid >> 6) ? wave0_total : 0);
  const int total = wave0_total + off_s[kBlock - 2];
  __syncthreads();
  off_s[tid] = off_i;
  nlj_s[tid] = nlj;
  if (tid == 0) off_s[kBlock] = total;
  // the first pair round's owner map, list entries and neighbour positions are requested now,
  // so that their latency overlaps the bonded terms
  constexpr int kPre = kChunk / kBlock < 2 ? kChunk / kBlock : 2;  // entry slots prefetched per thread
  const int c1_first = min(total, kChunk);
  if (active)
    for (int e = off_i; e < min(off_i + n_i, c1_first); e++) owner_s[e] = (unsigned char)tid;
  __syncthreads();
  static_assert(kChunk % kBlock == 0 && kPre >= 1, "tile of a multiple of the workgroup size");
  // list indices now, the neighbour positions once the hinges are done (the index loads are
  // not waited for before the bonded terms start)
  int pj0 = -1, pj1 = -1;
#pragma unroll
  for (int u = 0; u < kPre; u++) {
    const int e = tid + u * kBlock;
    if (e < c1_first) {
      const int i = owner_s[e];
      const int k = e - off_s[i];
      const int nl = nlj_s[i];
      const int j = (k < nl) ? d.lj_list[(size_t)k * d.Nb + s * np + i]
                             : d.q_list[(size_t)(k - nl) * d.Nq + (s * nq + tb.qslot[i])];
      if (u == 0) pj0 = j; else pj1 = j;
    }
  }
  // bending hinges -> LDS (the rows of a thread's first two hinges were loaded with the tables)
  auto hinge_loop = [&](auto wrap) {
    for (int h = tid; h < nh; h += kBlock) {
      const int4 hg = (h == tid) ? hg0 : (h == tid + kBlock) ? hg1 : reinterpret_cast<const int4*>(d.topo.hinge)[h];
      double hf[12];
      hinge_forces<decltype(wrap)::value>(own, hg, box, ibox, kb, hf);
#pragma unroll
      for (int k = 0; k < 9; k++) hc[9 * h + k] = hf[3 + k];
    }
  };
  if (bond_wrap) hinge_loop(Flag<true>{}); else hinge_loop(Flag<false>{});
  double4 pre0 = make_double4(0, 0, 0, 0), pre1 = pre0;  // no indexed array
  if (pj0 >= 0) pre0 = d.pos[pj0];
  if (kPre > 1 && pj1 >= 0) pre1 = d.pos[pj1];
  // stretching (src/bead.cpp:17-23) over the bead's edge row
  double sx = 0, sy = 0, sz = 0;
  double4 pi = make_double4(0, 0, 0, 0);
  auto stretch_loop = [&](auto wrap) {
    constexpr bool W = decltype(wrap)::value;
#pragma unroll
    for (int k = 0; k < kMaxDeg; k++) {
      if (k >= ecnt) break;
      const int o = ((k < 4 ? erow.x : erow.y) >> (8 * (k & 3))) & 0xff;
      const double l0 = (k & 1) ? len[k >> 1].y : len[k >> 1].x;
      const double4 po = own[o];
      const double dx = near_diff<W>(pi.x - po.x, box, ibox), dy = near_diff<W>(pi.y - po.y, box, ibox),
                   dz = near_diff<W>(pi.z - po.z, box, ibox);
      const double L2 = dx * dx + dy * dy + dz * dz, iL = rsqrt(L2), L = L2 * iL;
      const double c = (l0 - L) * ks * iL;
      sx += dx * c; sy += dy * c; sz += dz * c;
    }
  };
  if (active) {
    pi = own[tid];
    if (bond_wrap) stretch_loop(Flag<true>{}); else stretch_loop(Flag<false>{});
  }
  // bending gather (fixed hinge order); hinge bead 0 takes minus the sum of the other three,
  // formed as the hinge computes it
  __syncthreads();
  double bx = 0, by = 0, bz = 0;
  if (active) {
    const int hw[8] = {hrow0.x, hrow0.y, hrow0.z, hrow0.w, hrow1.x, hrow1.y, hrow1.z, hrow1.w};
#pragma unroll
    for (int k = 0; k < kMaxHinc; k++) {
      if (k >= hcnt) break;
      const int sl = (hw[k >> 1] >> (16 * (k & 1))) & 0xffff;
      const double* q = hc + 9 * (sl >> 2);
      if ((sl & 3) == 0) {
        bx += -(q[0] + q[3] + q[6]); by += -(q[1] + q[4] + q[7]); bz += -(q[2] + q[5] + q[8]);
      } else {
        const double* f = q + 3 * ((sl & 3) - 1);
        bx += f[0]; by += f[1]; bz += f[2];
      }
    }
  }
  // pair terms: all threads evaluate the subunit's list entries kChunk at a time; each bead
  // then adds its own entries in list order, the sum of a sequential per-bead loop
  PT lx = 0, ly = 0, lz = 0, ex = 0, ey = 0, ez = 0;
  {
    PairCtx<PT> ctx;
    ctx.owner = owner_s; ctx.off = off_s; ctx.nlj = nlj_s; ctx.tb = &tb; ctx.own = own; ctx.qq = qq_s; ctx.yk = yk_s;
    ctx.centres = d.centres; ctx.cs = cs; ctx.box = box; ctx.hbox = hbox; ctx.ibox = ibox; ctx.far_rf2 = P.far_rf2;
    ctx.ecut2 = (PT)P.ecut2; ctx.elj = (PT)P.elj_att; ctx.kappa = (PT)P.kappa; ctx.ecut_el2 = (PT)P.ecut_el2;
    ctx.nq = nq; ctx.split = P.split;
    auto entry = [&](int le, int c0, double4 pj) {
      if (pair_wrap) pair_entry<PT, kChunk, true>(le, c0, pj, ctx, pf);
      else pair_entry<PT, kChunk, false>(le, c0, pj, ctx, pf);
    };
    for (int c0 = 0; c0 < total; c0 += kChunk) {
      __syncthreads();
      const int c1 = min(total, c0 + kChunk);
      if (active && c0 > 0)  // the first round's map was made before the bonded terms
        for (int e = max(off_i, c0); e < min(off_i + n_i, c1); e++) owner_s[e - c0] = (unsigned char)tid;
      __syncthreads();
      int e_next = c0 + tid;
      if (c0 == 0) {  // the prefetched positions, one call per entry slot (no indexed array)
        if (tid < c1) entry(tid, 0, pre0);
        if (kPre > 1 && tid + kBlock < c1) entry(tid + kBlock, 0, pre1);
        e_next = tid + kPre * kBlock;
      }
      {
        for (int e = e_next; e < c1; e += kBlock) {
          const int i = owner_s[e - c0];
          const int k = e - off_s[i];
          const int nl = nlj_s[i];
          const int j = (k < nl) ? d.lj_list[(size_t)k * d.Nb + s * np + i]
                                 : d.q_list[(size_t)(k - nl) * d.Nq + (s * nq + tb.qslot[i])];
          entry(e - c0, c0, d.pos[j]);
        }
      }
      __syncthreads();
      if (active) {
        const int e1 = min(off_i + n_i, c1);
        for (int e = max(off_i, c0); e < e1; e++) {
          const PT fx = pf[e - c0], fy = pf[kChunk + e - c0], fz = pf[2 * kChunk + e - c0];
          if (e - off_i < nlj) { lx += fx; ly += fy; lz += fz; }
          else { ex += fx; ey += fy; ez += fz; }
        }
      }
    }
  }
  __syncthreads();

  double part[kPartStride];
#pragma unroll
  for (int k = 0; k < kPartStride; k++) part[k] = 0.0;
  if (active) {
    // src/forces.cpp:141-144: sforce + bforce + (eForce + ljForce); in split mode eForce is
    // the near-field list sum plus the far field of the charged bead
    double exd = (double)ex, eyd = (double)ey, ezd = (double)ez;
    if (d.split && tb.qslot[tid] >= 0) {
      const double* qf = d.qfar + 3 * ((size_t)s * nq + tb.qslot[tid]);
      exd = exd + qf[0]; eyd = eyd + qf[1]; ezd = ezd + qf[2];
    }
    const double elx = exd + (double)lx, ely = eyd + (double)ly, elz = ezd + (double)lz;
    double Fx = sx + bx + elx, Fy = sy + by + ely, Fz = sz + bz + elz;
    if (d.nbig) {  // src/forces.cpp:234-235: tforce + ljforce + eforce of the big-bead terms
      double bl[3] = {0, 0, 0}, be[3] = {0, 0, 0};
      big_bead_terms<false>(d, P, own[tid], tid, bl, be, nullptr, nullptr);
      Fx = (Fx + bl[0]) + be[0]; Fy = (Fy + bl[1]) + be[1]; Fz = (Fz + bl[2]) + be[2];
    }
    if (MODE == MODE_FORCES) {
      if (d.out_f) { d.out_f[3 * (size_t)gi] = Fx; d.out_f[3 * (size_t)gi + 1] = Fy; d.out_f[3 * (size_t)gi + 2] = Fz; }
      if (d.out_s) { d.out_s[3 * (size_t)gi] = sx; d.out_s[3 * (size_t)gi + 1] = sy; d.out_s[3 * (size_t)gi + 2] = sz; }
      if (d.out_b) { d.out_b[3 * (size_t)gi] = bx; d.out_b[3 * (size_t)gi + 1] = by; d.out_b[3 * (size_t)gi + 2] = bz; }
      if (d.out_e) { d.out_e[3 * (size_t)gi] = exd; d.out_e[3 * (size_t)gi + 1] = eyd; d.out_e[3 * (size_t)gi + 2] = ezd; }
      if (d.out_lj) { d.out_lj[3 * (size_t)gi] = lx; d.out_lj[3 * (size_t)gi + 1] = ly; d.out_lj[3 * (size_t)gi + 2] = lz; }
      if (init) { d.fx[gi] = Fx; d.fy[gi] = Fy; d.fz[gi] = Fz; }
    } else if (MODE == MODE_MD) {
      // second half kick (src/md.cpp:451-455)
      const double e = S.expfac, c = S.kick_c / m;
      vx = vx * e + Fx * c; vy = vy * e + Fy * c; vz = vz * e + Fz * c;
      st_stream(&d.vx[gi], vx); st_stream(&d.vy[gi], vy); st_stream(&d.vz[gi], vz);
      st_stream(&d.fx[gi], Fx); st_stream(&d.fy[gi], Fy); st_stream(&d.fz[gi], Fz);
    } else {
      // BD: tforce = F + (fran - mean) + fdrag; v += tforce dt/(2m) (src/md.cpp:462-498)
      double ux, uy, uz;
      bd_draws_skip(S.rng, skip, ux, uy, uz);
      const double amp = P.bd_amp[tid];
      const double rx = amp * ux - S.mean_fran[0], ry = amp * uy - S.mean_fran[1], rz = amp * uz - S.mean_fran[2];
      Fx = Fx + rx + d.gx[gi]; Fy = Fy + ry + d.gy[gi]; Fz = Fz + rz + d.gz[gi];
      const double c = 0.5 * dt / m;
      vx = vx + Fx * c; vy = vy + Fy * c; vz = vz + Fz * c;
      st_stream(&d.vx[gi], vx); st_stream(&d.vy[gi], vy); st_stream(&d.vz[gi], vz);
      st_stream(&d.fx[gi], Fx); st_stream(&d.fy[gi], Fy); st_stream(&d.fz[gi], Fz);
    }
    part[0] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    part[1] = vx; part[2] = vy; part[3] = vz;
    if (P.brownian) {
      // noise of the next step (or, at a start, of the first step: the engine state the start was
      // given, the seed for a fresh start) for its mean
      const uint32_t nxt = init ? S.rng : lcg_mulmod(S.rng, P.mult6n);
      double ux, uy, uz;
      bd_draws_skip(nxt, skip, ux, uy, uz);
      const double amp = P.bd_amp[tid];
      part[4] = amp * ux; part[5] = amp * uy; part[6] = amp * uz;
    }
  }
  // per-subunit sums, fixed shape: DPP wave sums then the two waves in order (slots 7-8 are
  // the list sizes written by k_rebuild; the BD noise slots 4-6 are not used by NH MD)
  constexpr int kSums = MODE == MODE_MD ? 4 : 7;
#pragma unroll
  for (int k = 0; k < kSums; k++) {
    const double w = wave_total(part[k]);
    if ((tid & 63) == 0) red[tid >> 6][k] = w;
  }
  __syncthreads();
  if (tid < kSums) d.part[tid * (size_t)d.Ns + s] = red[0][tid] + red[1][tid];
}

// One workgroup per replica: fixed-order sums of the per-subunit partials, then the scalar
// bookkeeping of src/md.cpp:400-406, 456-460, 469-478, 506-514.
constexpr int kFinThreads = 512;

__global__ __launch_bounds__(kFinThreads) void k_finalize(Dev d, int init) {
  __shared__ double wred[kFinThreads / 64][kPartSlots];
  __shared__ double tot[kPartSlots];
  __shared__ double cQ[kMaxChain], cinvQ[kMaxChain], cdof[kMaxChain], cxi[kMaxChain], ceta[kMaxChain];
  const int r = blockIdx.x, tid = threadIdx.x;
  const RepParams& P = d.rp[r];
  RepState& S = d.rs[r];
  // everything the serial tail reads is requested now, so its latency overlaps the sums: the
  // chain constants and state to LDS (lane j: bath j), the scalars to registers
  const int q = P.q_chain, brownian = P.brownian;
  const bool lists = init || S.rebuild;
  // a (re)start continues from the chain state in RepState (zero for a fresh start) and, when
  // resuming, from the kinetic energy the interrupted run's chain would have used
  const int resume = init ? S.resume : 0;
  const double ke_saved = S.ke;
  double xi_l = 0.0;
  if (tid < kMaxChain) {
    cQ[tid] = P.Qb[tid]; cinvQ[tid] = P.inv_Qb[tid]; cdof[tid] = P.dof[tid];
    const double xn = init ? S.xi[tid] : S.xi_n[tid];
    cxi[tid] = xn; ceta[tid] = init ? S.eta[tid] : S.eta_n[tid];
    if (!init && !brownian) xi_l = xn;
  }
  const double dt = P.dt, T = P.T, inv_nb = P.inv_nb;
  const uint32_t mult6n = P.mult6n, rng0 = S.rng;
  const long long step0 = S.step, rebuilds0 = S.rebuilds;
  // slot-major partials: a wave reads 64 consecutive subunits of one slot per load; each thread
  // adds its subunits tid, tid + kFinThreads, ... in that order, with the loads of kFinBatch of
  // them in flight together. Slots 4-6 (BD noise) and 7-8 (list sizes, written at list builds)
  // are read and reduced only when they are used.
  constexpr int kFinBatch = 4;
  double acc[kPartSlots];
#pragma unroll
  for (int k = 0; k < kPartSlots; k++) acc[k] = 0.0;
  const int s_end = P.s0 + P.ns;
  const size_t ns_all = d.Ns;
  for (int s0 = P.s0 + tid; s0 < s_end; s0 += kFinThreads * kFinBatch) {
    double v[kFinBatch][kPartSlots];
#pragma unroll
    for (int i = 0; i < kFinBatch; i++) {
      const int s = s0 + i * kFinThreads;
#pragma unroll
      for (int k = 0; k < kPartSlots; k++) {
        const bool used = s < s_end && (k < 4 || (k < 7 ? brownian != 0 : lists));
        v[i][k] = used ? d.part[k * ns_all + s] : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < kFinBatch; i++)
#pragma unroll
      for (int k = 0; k < kPartSlots; k++) acc[k] += v[i][k];
  }
#pragma unroll
  for (int k = 0; k < kPartSlots; k++) {
    const bool used = k < 4 || (k < 7 ? brownian != 0 : lists);  // uniform over the workgroup
    const double w = used ? wave_total(acc[k]) : 0.0;
    if ((tid & 63) == 0) wred[tid >> 6][k] = w;
  }
  __syncthreads();
  if (tid < kPartSlots) {
    double t = 0.0;
    for (int w = 0; w < kFinThreads / 64; w++) t += wred[w][tid];
    tot[tid] = t;
  }
  __syncthreads();
  if (tid >= 64) return;
  const double ke = resume ? ke_saved : tot[0];
  if (!brownian) {
    // forward chain of this step (src/functions.cpp:33-58 in bath order): bath j's update
    // reads xi[j+1] before that bath is updated, so lane j + 1 evaluates both exponentials of
    // bath j from its own xi at once
    const double e1 = exp(-0.5 * dt * xi_l), e2 = exp(-0.25 * dt * xi_l);
    double E1[kMaxChain], E2[kMaxChain];
#pragma unroll
    for (int j = 0; j < kMaxChain; j++) {
      E1[j] = __shfl(e1, j + 1, 64);
      E2[j] = __shfl(e2, j + 1, 64);
    }
    if (tid != 0) return;
    // bath constants are read from LDS where used; the chain state is in registers
    const double* Qb = cQ;
    const double* invQ = cinvQ;
    const double* dof = cdof;
    double xi[kMaxChain + 1], eta[kMaxChain];
#pragma unroll
    for (int j = 0; j < kMaxChain; j++) {
      xi[j] = cxi[j];
      eta[j] = ceta[j];
    }
    xi[kMaxChain] = 0.0;
    if (!init) {
      // end of this step: eta update, forward chain with the new kinetic energy
#pragma unroll
      for (int j = 0; j < kMaxChain; j++)
        if (j < q && Qb[j] != 0) eta[j] = eta[j] + 0.5 * dt * xi[j];
#pragma unroll
      for (int j = 0; j < kMaxChain; j++) {
        if (j >= q || Qb[j] == 0) continue;
        const double g = j == 0 ? 2 * ke - dof[0] * 1.0 * T : Qb[j - 1] * xi[j - 1] * xi[j - 1] - dof[j] * 1.0 * T;
        xi[j] = xi[j] * E1[j] + 0.5 * dt * invQ[j] * g * E2[j];
      }
    }
#pragma unroll
    for (int j = 0; j < kMaxChain; j++) { S.xi[j] = xi[j]; S.eta[j] = eta[j]; }
    // first thermostat half of the next step: reverse chain, eta update
#pragma unroll
    for (int j = kMaxChain - 1; j >= 0; j--)
      if (j < q) chain_xi(j, xi, Qb, invQ, dof, T, dt, ke);
#pragma unroll
    for (int j = 0; j < kMaxChain; j++)
      if (j < q && Qb[j] != 0) eta[j] = eta[j] + 0.5 * dt * xi[j];
#pragma unroll
    for (int j = 0; j < kMaxChain; j++) { S.xi_n[j] = xi[j]; S.eta_n[j] = eta[j]; }
    const double ef = exp(-0.5 * dt * xi[0]);
    S.expfac = ef;
    S.kick_c = 0.5 * dt * sqrt(ef);
  }
  if (tid != 0) return;
  S.ke = ke;
  // non-finite guard: the kinetic energy sums every velocity, so one non-finite bead shows here
  const double vsum = tot[0] + tot[1] + tot[2] + tot[3];
  if (!isfinite(vsum) && !S.nonfinite) {
    S.nonfinite = 1;
    S.bad_step = init ? step0 : step0 + 1;
  }
  S.resume = 0;
  if (init) {
    S.mean_v[0] = S.mean_v[1] = S.mean_v[2] = 0.0;
    S.vcom = 0.0;
  } else {
    const double mx = tot[1] * inv_nb, my = tot[2] * inv_nb, mz = tot[3] * inv_nb;
    S.mean_v[0] = mx; S.mean_v[1] = my; S.mean_v[2] = mz;
    S.vcom = sqrt(mx * mx + my * my + mz * mz);
    S.step = step0 + 1;
  }
  if (brownian) {
    S.rng = init ? rng0 : lcg_mulmod(rng0, mult6n);  // a (re)start keeps the engine state it was given
    S.mean_fran[0] = tot[4] * inv_nb; S.mean_fran[1] = tot[5] * inv_nb; S.mean_fran[2] = tot[6] * inv_nb;
  }
  if (lists) {
    if (!init) S.rebuilds = rebuilds0 + 1;
    S.lj_entries = (long long)tot[7];
    S.q_entries = (long long)tot[8];
  }
  S.rebuild = 0;
}

__global__ void k_apply_com(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  const int s = b / d.topo.np;
  const RepState& S = d.rs[rep_of(d, s)];
  d.vx[b] -= S.mean_v[0]; d.vy[b] -= S.mean_v[1]; d.vz[b] -= S.mean_v[2];
}
__global__ void k_clear_com(Dev d) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.R) return;
  d.rs[r].mean_v[0] = d.rs[r].mean_v[1] = d.rs[r].mean_v[2] = 0.0;
}

// Energies (src/md.cpp:553-593): per bead se, be, ne, ce, ke. Pair energies go to the lower
// bead id of each pair (src/energies.cpp:31-92), using the current lists.
// ------------------------------------------------------------------------------------------
// Far field of the uncut Yukawa (split mode): for subunit s, the Yukawa force on each of its
// charged beads from the charged beads of every subunit t of its replica whose centre bead
// lies beyond far_rf (minimum image of the centre beads), pair terms as src/forces.cpp:104-107
// with minimum-image bead vectors. Threads take subunits t in a fixed stride, then a
// fixed-shape reduction sums them: deterministic. With `energy`, also the unshifted pair
// energies of pairs with t > s (the reference gives a pair's energy to its lower id,
// src/energies.cpp:24-58).
constexpr int kFarQ = 8;         // charged beads per template the far field supports
constexpr int kFarThreads = 128;

__global__ __launch_bounds__(kFarThreads) void k_far_exact(Dev d, int energy) {
  __shared__ double red[kFarThreads / 64][kFarQ * 4];
  __shared__ double qq_s[kFarQ * kFarQ], yk_s[kFarQ * kFarQ];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = rep_of(d, s);
  const RepParams& P = d.rp[r];
  const int nq = d.topo.nq;
  const double box = P.box, hbox = P.hbox, kappa = P.kappa;
  for (int k = tid; k < nq * nq; k += kFarThreads) { qq_s[k] = P.qq[k]; yk_s[k] = P.yk[k]; }
  const double4 ca = d.centres[s];
  double4 qa[kFarQ];
#pragma unroll
  for (int a = 0; a < kFarQ; a++) qa[a] = a < nq ? d.qpos[(size_t)s * nq + a] : make_double4(0, 0, 0, 0);
  double f[kFarQ][3], e[kFarQ];
#pragma unroll
  for (int a = 0; a < kFarQ; a++) { f[a][0] = f[a][1] = f[a][2] = 0.0; e[a] = 0.0; }
  __syncthreads();
  for (int t = P.s0 + tid; t < P.s0 + P.ns; t += kFarThreads) {
    const double4 cb = d.centres[t];
    const double cx = min_img(ca.x - cb.x, box, hbox), cy = min_img(ca.y - cb.y, box, hbox),
                 cz = min_img(ca.z - cb.z, box, hbox);
    if (t == s || !(cx * cx + cy * cy + cz * cz > P.far_rf2)) continue;
    const bool count_e = energy && t > s;
    for (int b = 0; b < nq; b++) {
      const double4 pb = d.qpos[(size_t)t * nq + b];
#pragma unroll
      for (int a = 0; a < kFarQ; a++) {
        if (a >= nq) break;
        const double dx = min_img(qa[a].x - pb.x, box, hbox), dy = min_img(qa[a].y - pb.y, box, hbox),
                     dz = min_img(qa[a].z - pb.z, box, hbox);
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rr = sqrt(r2);
        const double qq = qq_s[a * nq + b], yk = yk_s[a * nq + b];
        const double ex = exp(-kappa * rr);
        const double c = (qq * ex / r2) * (kappa + 1 / rr) * yk;
        f[a][0] += dx * c; f[a][1] += dy * c; f[a][2] += dz * c;
        if (count_e) e[a] += yk * ((qq * ex) / rr);
      }
    }
  }
  // fixed-shape reduction: wave sums, then the two waves in order
#pragma unroll
  for (int a = 0; a < kFarQ; a++) {
    if (a >= nq) break;
    const double w0 = wave_sum(f[a][0]), w1 = wave_sum(f[a][1]), w2 = wave_sum(f[a][2]), w3 = wave_sum(e[a]);
    if ((tid & 63) == 0) {
      red[tid >> 6][4 * a] = w0; red[tid >> 6][4 * a + 1] = w1; red[tid >> 6][4 * a + 2] = w2;
      red[tid >> 6][4 * a + 3] = w3;
    }
  }
  __syncthreads();
  if (tid < nq * 3) {
    const int a = tid / 3, c = tid - 3 * a;
    d.qfar[3 * ((size_t)s * nq + a) + c] = red[0][4 * a + c] + red[1][4 * a + c];
  }
  if (energy && tid < nq) d.efar[(size_t)s * nq + tid] = red[0][4 * tid + 3] + red[1][4 * tid + 3];
}

__global__ __launch_bounds__(kBlock) void k_energy(Dev d) {
  extern __shared__ double smem[];
  __shared__ double red[2][5];
  const int np = d.topo.np, nh = d.topo.nh, nq = d.topo.nq;
  double4* own = reinterpret_cast<double4*>(smem);
  double* he = reinterpret_cast<double*>(own + np);
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = rep_of(d, s);
  const RepParams& P = d.rp[r];
  const double box = P.box, hbox = P.hbox;
  const double4 cs = d.centres[s];
  const bool active = tid < np;
  const int gi = s * np + tid;
  if (active) own[tid] = d.pos[gi];
  __syncthreads();
  for (int h = tid; h < nh; h += kBlock) {
    double tmp[12];
    const int4 hg = reinterpret_cast<const int4*>(d.topo.hinge)[h];
    const double c = hinge_forces<true>(own, hg, box, P.ibox, P.kb, tmp);
    he[h] = P.kb * 0.25 * (1 - fabs(c));  // src/edge.cpp:64-66 after the flip: n0.n1 = |cos|
  }
  __syncthreads();
  double e[5] = {0, 0, 0, 0, 0};
  if (active) {
    const double4 pi = own[tid];
    const int ecnt = d.topo.e_cnt[tid];
    for (int k = 0; k < ecnt; k++) {  // src/bead.cpp:9-14
      const double4 po = own[d.topo.e_other[tid * kMaxDeg + k]];
      const double dx = min_img(pi.x - po.x, box, hbox), dy = min_img(pi.y - po.y, box, hbox),
                   dz = min_img(pi.z - po.z, box, hbox);
      const double L = sqrt(dx * dx + dy * dy + dz * dz);
      const double dl = d.topo.e_len0[tid * kMaxDeg + k] - L;
      e[0] += 0.25 * P.ks * dl * dl;
    }
    const int hcnt = d.topo.h_cnt[tid];
    for (int k = 0; k < hcnt; k++) e[1] += he[d.topo.h_slot[tid * kMaxHinc + k] >> 2];
    // LJ / WCA energies (src/energies.cpp:63-90): s = (s_i+s_j)/2, shifted attractive form
    const int ti = d.topo.type[tid];
    const unsigned long long mask_i = d.topo.att_mask[ti];
    const double si = d.topo.sigma[tid];
    const double ecut3 = P.ecut * P.ecut * P.ecut, ecut6 = ecut3 * ecut3;
    const int cl = min(d.lj_cnt[gi], d.lj_cap);
    for (int k = 0; k < cl; k++) {
      const int j = d.lj_list[(size_t)k * d.Nb + gi];
      if (j <= gi) continue;
      const double4 pj = d.pos[j];
      const int lj = tag_loc(pj.w);
      const double dx = min_img(pi.x - pj.x, box, hbox), dy = min_img(pi.y - pj.y, box, hbox),
                   dz = min_img(pi.z - pj.z, box, hbox);
      const double rr = sqrt(dx * dx + dy * dy + dz * dz);
      const double shc = (si + d.topo.sigma[lj]) / 2;
      if (rr >= P.ecut && rr >= kWcaFactor * shc) continue;
      const double s3 = shc * shc * shc, s6 = s3 * s3, r3 = rr * rr * rr, r6 = r3 * r3;
      const bool att = (mask_i >> d.topo.type[lj]) & 1ull;
      if (rr < kWcaFactor * shc && !att) {
        e[2] += (4 * 1.0 * (s6 / r6) * ((s6 / r6) - 1)) + 1.0;
      } else if (rr < P.ecut && att) {
        e[2] += (4 * P.elj_att * (s6 / r6) * ((s6 / r6) - 1)) - (4 * P.elj_att * ((s6 / ecut6) * ((s6 / ecut6) - 1)));
      }
    }
    // Yukawa energy shifted at ecut_el (src/energies.cpp:54-58)
    const int qi = d.topo.qslot[tid];
    if (qi >= 0) {
      const int gq = s * nq + qi;
      const int cq = min(d.q_cnt[gq], d.q_cap);
      for (int k = 0; k < cq; k++) {
        const int j = d.q_list[(size_t)k * d.Nq + gq];
        if (j <= gi) continue;
        const double4 pj = d.pos[j];
        const int qj = d.topo.qslot[tag_loc(pj.w)];
        const double dx = min_img(pi.x - pj.x, box, hbox), dy = min_img(pi.y - pj.y, box, hbox),
                     dz = min_img(pi.z - pj.z, box, hbox);
        const double rr = sqrt(dx * dx + dy * dy + dz * dz);
        const double qq = P.qq[qi * nq + qj];
        if (P.split) {  // uncut: unshifted (src/energies.cpp:54-58 with ecut_el = 0), near field only
          const double4 cb = d.centres[tag_sub(pj.w)];
          const double cx = min_img(cs.x - cb.x, box, hbox), cy = min_img(cs.y - cb.y, box, hbox),
                       cz = min_img(cs.z - cb.z, box, hbox);
          if (cx * cx + cy * cy + cz * cz <= P.far_rf2) e[3] += P.yk[qi * nq + qj] * ((qq * exp(-P.kappa * rr)) / rr);
        } else if (rr < P.ecut_el) {
          e[3] += P.yk[qi * nq + qj] * ((qq * exp(-P.kappa * rr)) / rr - (qq * exp(-P.kappa * P.ecut_el)) / P.ecut_el);
        }
      }
      if (P.split) e[3] += d.efar[gq];
    }
    if (d.nbig) {  // src/energies.cpp:129-134: the big-bead terms at full value
      e[2] += d.big_f[2 * (size_t)gi].w;
      e[3] += d.big_f[2 * (size_t)gi + 1].w;
    }
    const double m = d.topo.mass[tid];
    const double vx = d.vx[gi], vy = d.vy[gi], vz = d.vz[gi];
    e[4] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    if (d.ebead)
      for (int k = 0; k < 5; k++) d.ebead[5 * (size_t)gi + k] = e[k];
  }
  // energy order of the output: se, be, ne, ce, ke
  for (int k = 0; k < 5; k++) {
    const double w = wave_sum(e[k]);
    if ((tid & 63) == 0) red[tid >> 6][k] = w;
  }
  __syncthreads();
  if (tid < 5) d.epart[(size_t)s * 5 + tid] = red[0][tid] + red[1][tid];
}

// Big-bead terms of every subunit bead with their energies, for the energy entries: one thread
// per bead; out[2b] = (ljforce, ne), out[2b + 1] = (eforce, ce). The step and force paths
// evaluate the forces inside k_force instead.
__global__ __launch_bounds__(256) void k_big(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  const double4 p = d.pos[b];
  double lj[3] = {0, 0, 0}, el[3] = {0, 0, 0}, ne = 0, ce = 0;
  big_bead_terms<true>(d, d.rp[0], p, tag_loc(p.w), lj, el, &ne, &ce);
  d.big_f[2 * (size_t)b] = make_double4(lj[0], lj[1], lj[2], ne);
  d.big_f[2 * (size_t)b + 1] = make_double4(el[0], el[1], el[2], ce);
}

__global__ void k_bd_uniforms(Dev d, uint32_t state, int nb, double* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double ux, uy, uz;
  bd_draws(state, b, ux, uy, uz);
  out[3 * (size_t)b] = ux; out[3 * (size_t)b + 1] = uy; out[3 * (size_t)b + 2] = uz;
}

// ------------------------------------------------------------------------------------------
static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return SF_E_HIP;
  }
  return SF_OK;
}

int launch_centres(const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_centres, dim3((d.Ns + 255) / 256), dim3(256), 0, st, d);
  return check_launch();
}

int launch_integrate(const Dev& d, int mode, hipStream_t st) {
  dim3 g((d.Nb + 255) / 256), b(256);
  if (mode == MODE_MD) hipLaunchKernelGGL(k_integrate<MODE_MD>, g, b, 0, st, d);
  else hipLaunchKernelGGL(k_integrate<MODE_BD>, g, b, 0, st, d);
  return check_launch();
}

int launch_moved(const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_moved, dim3((d.Nb + 255) / 256), dim3(256), 0, st, d);
  return check_launch();
}
int launch_lists_done(const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_lists_done, dim3((d.R + 63) / 64), dim3(64), 0, st, d);
  return check_launch();
}
int launch_rebuild(const Dev& d, int force_rebuild, hipStream_t st) {
  if (!d.cells) {
    hipLaunchKernelGGL(k_rebuild, dim3(d.Ns), dim3(kBlock), 0, st, d, force_rebuild);
    return check_launch();
  }
  if (force_rebuild) {  // outside the step loop no integrate kernel has cleared the cells
    hipLaunchKernelGGL(k_cells_clear, dim3(256), dim3(256), 0, st, d);
    if (int rc = check_launch()) return rc;
  }
  hipLaunchKernelGGL(k_cell_fill, dim3(d.Ns), dim3(kBlock), 0, st, d, force_rebuild);
  if (int rc = check_launch()) return rc;
  hipLaunchKernelGGL(k_cell_query, dim3(d.Ns), dim3(kBlock), 0, st, d, force_rebuild);
  return check_launch();
}

bool tile_supported(int tile) { return tile == 128 || tile == 256 || tile == 384 || tile == 512; }

template <int CH>
static void launch_force_tile(const Dev& d, int mode, int init, int fp32, size_t sm, hipStream_t st) {
  dim3 g(d.Ns), b(kBlock);
  if (fp32) {
    if (mode == MODE_FORCES) hipLaunchKernelGGL((k_force<MODE_FORCES, float, CH>), g, b, sm, st, d, init);
    else if (mode == MODE_MD) hipLaunchKernelGGL((k_force<MODE_MD, float, CH>), g, b, sm, st, d, init);
    else hipLaunchKernelGGL((k_force<MODE_BD, float, CH>), g, b, sm, st, d, init);
  } else {
    if (mode == MODE_FORCES) hipLaunchKernelGGL((k_force<MODE_FORCES, double, CH>), g, b, sm, st, d, init);
    else if (mode == MODE_MD) hipLaunchKernelGGL((k_force<MODE_MD, double, CH>), g, b, sm, st, d, init);
    else hipLaunchKernelGGL((k_force<MODE_BD, double, CH>), g, b, sm, st, d, init);
  }
}

int launch_force(const Dev& d, int mode, int init, int fp32, hipStream_t st) {
  const size_t sm = force_lds_bytes(d.topo.np, d.topo.nh, d.topo.nq, d.tile);
  if (d.tile == 128) launch_force_tile<128>(d, mode, init, fp32, sm, st);
  else if (d.tile == 512) launch_force_tile<512>(d, mode, init, fp32, sm, st);
  else if (d.tile == 256) launch_force_tile<256>(d, mode, init, fp32, sm, st);
  else if (d.tile == 384) launch_force_tile<384>(d, mode, init, fp32, sm, st);
  else { set_error("unsupported k_force tile"); return SF_E_ARG; }
  return check_launch();
}

__global__ void k_fill(double* p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int launch_fill(double* p, size_t n, double v, hipStream_t st) {
  if (n == 0) return SF_OK;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return check_launch();
}

int launch_far_exact(const Dev& d, int energy, hipStream_t st) {
  if (d.topo.nq > kFarQ) { set_error("the far field supports at most 8 charged beads per template"); return SF_E_ARG; }
  hipLaunchKernelGGL(k_far_exact, dim3(d.Ns), dim3(kFarThreads), 0, st, d, energy);
  return check_launch();
}

int launch_finalize(const Dev& d, int init, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(d.R), dim3(kFinThreads), 0, st, d, init);
  return check_launch();
}

int launch_apply_com(const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_com, dim3((d.Nb + 255) / 256), dim3(256), 0, st, d);
  int rc = check_launch();
  if (rc) return rc;
  hipLaunchKernelGGL(k_clear_com, dim3((d.R + 63) / 64), dim3(64), 0, st, d);
  return check_launch();
}

int launch_energy(const Dev& d, hipStream_t st) {
  size_t sm = (size_t)d.topo.np * sizeof(double4) + (size_t)d.topo.nh * sizeof(double);
  hipLaunchKernelGGL(k_energy, dim3(d.Ns), dim3(kBlock), sm, st, d);
  return check_launch();
}

int launch_big(const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_big, dim3((d.Nb + 255) / 256), dim3(256), 0, st, d);
  return check_launch();
}

int launch_bd_uniforms_state(uint32_t state, int nb, double* out_dev, const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_bd_uniforms, dim3((nb + 255) / 256), dim3(256), 0, st, d, state, nb, out_dev);
  return check_launch();
}

}  // namespace sf
