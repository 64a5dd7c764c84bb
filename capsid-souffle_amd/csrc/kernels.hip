// HIP kernels for gfx950 (MI355X): one SOUFFLE step is
//   k_integrate  one thread per bead: pending COM removal, first half kick (NH) or drag +
//                half kick (BD), drift, periodic wrap, list-displacement check
//   k_force      one 128-thread workgroup per subunit: (list rebuild when flagged) +
//                stretch + bending (hinge terms staged in LDS, gathered per bead through the
//                template's incidence CSR) + LJ/WCA and Yukawa gathers over per-bead lists +
//                second half kick (NH) or noise + drag + half kick (BD) + per-subunit sums
//   k_finalize   one workgroup per replica: fixed-order reductions, Nose-Hoover chain,
//                centre-of-mass velocity, BD engine advance
// No atomics touch forces or sums: each bead's force is gathered by one thread in a fixed
// order and every reduction has a fixed shape, so runs are bit-reproducible.
#include <hip/hip_runtime.h>

#include "sf_internal.hpp"

namespace sf {

__device__ __forceinline__ double min_img(double v, double box, double hbox) {
  // single conditional shift per axis, as src/forces.cpp:85-91 and src/functions.h:42-56
  if (v > hbox) v -= box;
  else if (v < -hbox) v += box;
  return v;
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

// The 3 uniform draws of bead b (replica-local) for the step whose engine state is `state`:
// draws 6b+1 .. 6b+6 of the stream, x, y, z order (src/md.cpp:464-466).
__device__ __forceinline__ void bd_draws(uint32_t state, int b, double& ux, double& uy, double& uz) {
  uint32_t x = lcg_mulmod(state, lcg_powmod(6ull * (unsigned long long)b));
  uint32_t g[6];
#pragma unroll
  for (int k = 0; k < 6; k++) { x = lcg_mulmod(x, kLcgMul); g[k] = x; }
  ux = canonical_uniform(g[0], g[1]);
  uy = canonical_uniform(g[2], g[3]);
  uz = canonical_uniform(g[4], g[5]);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Nose-Hoover chain variable j (src/functions.cpp:33-58).
__device__ void chain_xi(int j, double* xi, const double* Qb, const double* dof, double T, double dt,
                         double ke) {
  if (Qb[j] == 0) return;
  if (j != 0) {
    xi[j] = xi[j] * exp(-0.5 * dt * xi[j + 1]) +
            0.5 * dt * (1.0 / Qb[j]) * (Qb[j - 1] * xi[j - 1] * xi[j - 1] - dof[j] * 1.0 * T) *
                exp(-0.25 * dt * xi[j + 1]);
  } else {
    xi[j] = xi[j] * exp(-0.5 * dt * xi[j + 1]) +
            0.5 * dt * (1.0 / Qb[j]) * (2 * ke - dof[j] * 1.0 * T) * exp(-0.25 * dt * xi[j + 1]);
  }
}

// ------------------------------------------------------------------------------------------
__global__ void k_centres(Dev d) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.Ns) return;
  d.centres[s] = d.pos[(size_t)s * d.topo.np + d.topo.centre];
}

// First half of the step (src/md.cpp:400-431): NH: v = v e + F (dt/2 sqrt(e))/m;
// BD: fdrag = -m v / damp, v += F dt/(2m). Then x += v dt and one periodic wrap
// (src/bead.h:72-81). The centre-of-mass velocity left by the previous step is removed first.
template <int MODE>
__global__ __launch_bounds__(256) void k_integrate(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  double4 p = d.pos[b];
  const int s = tag_sub(p.w), l = tag_loc(p.w);
  const int r = d.sub_rep[s];
  const RepParams& P = d.rp[r];
  RepState& S = d.rs[r];
  const double m = d.topo.mass[l];
  double vx = d.vx[b] - S.mean_v[0], vy = d.vy[b] - S.mean_v[1], vz = d.vz[b] - S.mean_v[2];
  const double fx = d.fx[b], fy = d.fy[b], fz = d.fz[b];
  if (MODE == MODE_MD) {
    const double e = S.expfac, c = S.kick_c / m;
    vx = vx * e + fx * c; vy = vy * e + fy * c; vz = vz * e + fz * c;
  } else {
    const double cd = -1.0 * m / P.damp;
    d.gx[b] = vx * cd; d.gy[b] = vy * cd; d.gz[b] = vz * cd;
    const double c = 0.5 * P.dt / m;
    vx = vx + fx * c; vy = vy + fy * c; vz = vz + fz * c;
  }
  const double box = P.box, hbox = P.hbox;
  p.x = p.x + vx * P.dt; p.y = p.y + vy * P.dt; p.z = p.z + vz * P.dt;
  if (p.x > hbox) p.x = p.x - box; else if (p.x < -hbox) p.x = p.x + box;
  if (p.y > hbox) p.y = p.y - box; else if (p.y < -hbox) p.y = p.y + box;
  if (p.z > hbox) p.z = p.z - box; else if (p.z < -hbox) p.z = p.z + box;
  d.pos[b] = p;
  d.vx[b] = vx; d.vy[b] = vy; d.vz[b] = vz;
  const double4 q = d.xb[b];
  const double ddx = min_img(p.x - q.x, box, hbox), ddy = min_img(p.y - q.y, box, hbox),
               ddz = min_img(p.z - q.z, box, hbox);
  if (ddx * ddx + ddy * ddy + ddz * ddz > P.half_skin2) S.rebuild = 1;  // same value from any thread
  if (l == d.topo.centre) d.centres[s] = p;
}

// ------------------------------------------------------------------------------------------
// Neighbour-list build for the beads of subunit s. Candidate subunits are those whose centre
// beads lie within the bound (pair cutoff + skin + twice the centre-to-bead radius); they are
// taken in ascending id, and each one's beads are staged in LDS and tested against every bead
// of s, so each list comes out in ascending global bead id, the order in which the
// reference's update_pairlist appends (src/functions.cpp:93-110).
// LJ list: beads of other subunits within (cutoff of the pair + skin), the cutoff being ecut
// for an attractive (type_i, type_j) row of lj_a and 1.12246205 (s_i + s_j)/2 otherwise
// (src/forces.cpp:112-138). Yukawa list: charged beads j != i of any subunit (src/forces.cpp:79)
// within ecut_el + skin.
__device__ void build_lists(const Dev& d, const RepParams& P, RepState& S, int s, int tid,
                            const double4* own, double4* stage, int* cand, unsigned char* cflag,
                            int* wcnt, int* ncand_sh, int& cnt_lj, int& cnt_q) {
  const int np = d.topo.np, nq = d.topo.nq;
  const double box = P.box, hbox = P.hbox;
  const double4 cs = d.centres[s];
  if (tid == 0) *ncand_sh = 0;
  __syncthreads();
  for (int base = 0; base < P.ns; base += kBlock) {
    const int t = P.s0 + base + tid;
    int flag = 0;
    if (base + tid < P.ns) {
      const double4 ct = d.centres[t];
      const double dx = min_img(cs.x - ct.x, box, hbox), dy = min_img(cs.y - ct.y, box, hbox),
                   dz = min_img(cs.z - ct.z, box, hbox);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (t != s && d2 < P.sub_rc_lj2) flag |= 1;
      if (nq > 0 && d2 < P.sub_rc_q2) flag |= 2;
    }
    const unsigned long long bal = __ballot(flag != 0);
    const int lane = tid & 63, wave = tid >> 6;
    const int prefix = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wave] = __popcll(bal);
    __syncthreads();
    const int off = *ncand_sh + (wave ? wcnt[0] : 0);
    if (flag) {
      const int slot = off + prefix;
      if (slot < kMaxCand) { cand[slot] = t; cflag[slot] = (unsigned char)flag; }
    }
    __syncthreads();
    if (tid == 0) *ncand_sh += wcnt[0] + wcnt[1];
    __syncthreads();
  }
  int nc = *ncand_sh;
  if (nc > kMaxCand) {
    if (tid == 0) atomicMax(&S.need_cand, nc);
    nc = kMaxCand;
  }
  const bool active = tid < np;
  const int gi = s * np + tid;
  double4 pi = active ? own[tid] : make_double4(0, 0, 0, 0);
  const int ti = active ? d.topo.type[tid] : 0;
  const double si = active ? d.topo.sigma[tid] : 0.0;
  const unsigned long long mask_i = active ? d.topo.att_mask[ti] : 0ull;
  const int qi = active ? d.topo.qslot[tid] : -1;
  const int gq = qi >= 0 ? s * nq + qi : -1;
  const double ecut_l = P.ecut + P.skin;
  const double ecut_l2 = ecut_l * ecut_l;
  cnt_lj = 0;
  cnt_q = 0;
  for (int c = 0; c < nc; c++) {
    const int t = cand[c];
    const int f = cflag[c];
    __syncthreads();
    if (tid < np) stage[tid] = d.pos[(size_t)t * np + tid];
    __syncthreads();
    if (!active) continue;
    if (f & 1) {
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
    if ((f & 2) && qi >= 0) {
      for (int kq = 0; kq < nq; kq++) {
        const int k = d.topo.qlocal[kq];
        const int j = t * np + k;
        if (j == gi) continue;
        const double4 pj = stage[k];
        const double dx = min_img(pi.x - pj.x, box, hbox), dy = min_img(pi.y - pj.y, box, hbox),
                     dz = min_img(pi.z - pj.z, box, hbox);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < P.rc_q_list2) {
          if (cnt_q < d.q_cap) d.q_list[(size_t)cnt_q * d.Nq + gq] = j;
          cnt_q++;
        }
      }
    }
  }
  if (active) {
    d.lj_cnt[gi] = cnt_lj;
    if (gq >= 0) d.q_cnt[gq] = cnt_q;
    d.xb[gi] = pi;
    if (cnt_lj > d.lj_cap) atomicMax(&S.need_lj, cnt_lj);
    if (cnt_q > d.q_cap) atomicMax(&S.need_q, cnt_q);
    const double4 pc = own[d.topo.centre];
    const double dx = min_img(pi.x - pc.x, box, hbox), dy = min_img(pi.y - pc.y, box, hbox),
                 dz = min_img(pi.z - pc.z, box, hbox);
    if (dx * dx + dy * dy + dz * dz > P.r_bound2) S.deformed = 1;
  }
}

// Bending of one hinge: E = Kb (1 - |cos t|), cos t the cosine between the normals of the two
// faces; equal to the reference's (grad S - (S/a0 grad a0 + S/a1 grad a1)) Kb / (4 a0 a1) with
// S = 4 |A0.A1| (src/edge.cpp:69-188). With e = x1-x0, a = x2-x0, b = x3-x0, u = e x a,
// v = e x b, c = u.v/(|u||v|): dc/du = g = (v/|v| - c u/|u|)/|u|, dc/dv = h likewise, and
// dc/dx2 = g x e, dc/dx3 = h x e, dc/dx1 = a x g + b x h, dc/dx0 = -(sum).
// Writes the 4 bead forces (12 doubles) and returns cos t.
__device__ __forceinline__ double hinge_forces(const double4* own, const int* hg, double box, double hbox,
                                               double Kb, double* out) {
  const double4 p0 = own[hg[0]], p1 = own[hg[1]], p2 = own[hg[2]], p3 = own[hg[3]];
  const double ex = min_img(p1.x - p0.x, box, hbox), ey = min_img(p1.y - p0.y, box, hbox),
               ez = min_img(p1.z - p0.z, box, hbox);
  const double ax = min_img(p2.x - p0.x, box, hbox), ay = min_img(p2.y - p0.y, box, hbox),
               az = min_img(p2.z - p0.z, box, hbox);
  const double bx = min_img(p3.x - p0.x, box, hbox), by = min_img(p3.y - p0.y, box, hbox),
               bz = min_img(p3.z - p0.z, box, hbox);
  const double ux = ey * az - ez * ay, uy = ez * ax - ex * az, uz = ex * ay - ey * ax;
  const double vx = ey * bz - ez * by, vy = ez * bx - ex * bz, vz = ex * by - ey * bx;
  const double iu = 1.0 / sqrt(ux * ux + uy * uy + uz * uz), iv = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
  const double c = (ux * vx + uy * vy + uz * vz) * iu * iv;
  const double gx = (vx * iv - ux * (c * iu)) * iu, gy = (vy * iv - uy * (c * iu)) * iu,
               gz = (vz * iv - uz * (c * iu)) * iu;
  const double hx = (ux * iu - vx * (c * iv)) * iv, hy = (uy * iu - vy * (c * iv)) * iv,
               hz = (uz * iu - vz * (c * iv)) * iv;
  const double k = (c < 0) ? -Kb : Kb;
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

// LJ / WCA gather over bead i's list (src/forces.cpp:112-139): attractive rows use
// 48 eps (s_i^6/r^6)((s_i^6/r^6) - 1/2)/r^2 below ecut with eps = elj_att, or 50 when
// type_j == 7; other pairs the WCA form with eps = 1 and s = (s_i + s_j)/2 below
// 1.12246205 s.
template <typename PT>
__device__ __forceinline__ void lj_gather(const Dev& d, const RepParams& P, int gi, int cnt, double4 pi,
                                          double si, unsigned long long mask_i, PT& fx, PT& fy, PT& fz) {
  const double box = P.box, hbox = P.hbox;
  const PT ecut2 = (PT)P.ecut2, elj = (PT)P.elj_att;
  const PT si_ = (PT)si;
  const PT si6 = si_ * si_ * si_ * si_ * si_ * si_;
  const PT replj2 = (PT)(kWcaFactor * kWcaFactor);
  for (int k = 0; k < cnt; k++) {
    const int j = d.lj_list[(size_t)k * d.Nb + gi];
    const double4 pj = d.pos[j];
    const int lj = tag_loc(pj.w);
    const int tj = d.topo.type[lj];
    const PT sj = (PT)d.topo.sigma[lj];
    const PT dx = (PT)min_img(pi.x - pj.x, box, hbox), dy = (PT)min_img(pi.y - pj.y, box, hbox),
             dz = (PT)min_img(pi.z - pj.z, box, hbox);
    const PT r2 = dx * dx + dy * dy + dz * dz;
    const PT shc = (si_ + sj) / (PT)2;
    PT s6, eps;
    bool on;
    if ((mask_i >> tj) & 1ull) {
      on = r2 < ecut2;
      s6 = si6;
      eps = (tj == 7) ? (PT)50 : elj;
    } else {
      on = r2 < replj2 * shc * shc;
      s6 = shc * shc * shc * shc * shc * shc;
      eps = (PT)1;
    }
    if (on) {
      const PT r6 = r2 * r2 * r2;
      const PT sr6 = s6 / r6;
      const PT f = (PT)48 * eps * (sr6 * (sr6 - (PT)0.5)) * ((PT)1 / r2);
      fx += dx * f; fy += dy * f; fz += dz * f;
    }
  }
}

// Screened Coulomb gather over charged bead i's list (src/forces.cpp:100-108):
// q_i q_j lb exp(-k r)/r^2 (k + 1/r) Y_ij r_vec below ecut_el.
template <typename PT>
__device__ __forceinline__ void yk_gather(const Dev& d, const RepParams& P, int gq, int qi, int cnt, double4 pi,
                                          PT& fx, PT& fy, PT& fz) {
  const double box = P.box, hbox = P.hbox;
  const int nq = d.topo.nq;
  const PT kappa = (PT)P.kappa, ecut_el2 = (PT)P.ecut_el2;
  for (int k = 0; k < cnt; k++) {
    const int j = d.q_list[(size_t)k * d.Nq + gq];
    const double4 pj = d.pos[j];
    const int qj = d.topo.qslot[tag_loc(pj.w)];
    const PT dx = (PT)min_img(pi.x - pj.x, box, hbox), dy = (PT)min_img(pi.y - pj.y, box, hbox),
             dz = (PT)min_img(pi.z - pj.z, box, hbox);
    const PT r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < ecut_el2) {
      const PT r = sqrt(r2);
      const PT f = ((PT)P.qq[qi * nq + qj] * exp(-kappa * r) / r2) * (kappa + (PT)1 / r) * (PT)P.yk[qi * nq + qj];
      fx += dx * f; fy += dy * f; fz += dz * f;
    }
  }
}

template <int MODE, typename PT>
__global__ __launch_bounds__(kBlock) void k_force(Dev d, int init, int force_rebuild) {
  extern __shared__ double smem[];
  __shared__ int cand[kMaxCand];
  __shared__ unsigned char cflag[kMaxCand];
  __shared__ int wcnt[2];
  __shared__ int ncand_sh;
  __shared__ double red[2][kPartStride];
  const int np = d.topo.np, nh = d.topo.nh, nq = d.topo.nq;
  double4* own = reinterpret_cast<double4*>(smem);
  double4* stage = own + np;
  double* hc = reinterpret_cast<double*>(stage + np);

  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = d.sub_rep[s];
  const RepParams& P = d.rp[r];
  RepState& S = d.rs[r];
  const double box = P.box, hbox = P.hbox;
  const bool active = tid < np;
  const int gi = s * np + tid;
  if (active) own[tid] = d.pos[gi];
  const bool rebuild = force_rebuild || S.rebuild != 0;
  __syncthreads();

  int cnt_lj = 0, cnt_q = 0;
  const int qi = active ? d.topo.qslot[tid] : -1;
  const int gq = qi >= 0 ? s * nq + qi : -1;
  if (rebuild) {
    build_lists(d, P, S, s, tid, own, stage, cand, cflag, wcnt, &ncand_sh, cnt_lj, cnt_q);
  } else if (active) {
    cnt_lj = d.lj_cnt[gi];
    if (gq >= 0) cnt_q = d.q_cnt[gq];
  }
  // bending hinges -> LDS
  for (int h = tid; h < nh; h += kBlock) hinge_forces(own, d.topo.hinge + 4 * h, box, hbox, P.kb, hc + 12 * h);
  __syncthreads();

  double part[kPartStride];
#pragma unroll
  for (int k = 0; k < kPartStride; k++) part[k] = 0.0;
  if (active) {
    const double4 pi = own[tid];
    // stretching (src/bead.cpp:17-23)
    double sx = 0, sy = 0, sz = 0;
    for (int k = d.topo.e_off[tid]; k < d.topo.e_off[tid + 1]; k++) {
      const double4 po = own[d.topo.e_other[k]];
      const double dx = min_img(pi.x - po.x, box, hbox), dy = min_img(pi.y - po.y, box, hbox),
                   dz = min_img(pi.z - po.z, box, hbox);
      const double L = sqrt(dx * dx + dy * dy + dz * dz);
      const double c = (d.topo.e_len0[k] - L) * P.ks / L;
      sx += dx * c; sy += dy * c; sz += dz * c;
    }
    // bending gather (fixed hinge order)
    double bx = 0, by = 0, bz = 0;
    for (int k = d.topo.h_off[tid]; k < d.topo.h_off[tid + 1]; k++) {
      const int sl = d.topo.h_slot[k];
      const double* q = hc + 12 * (sl >> 2) + 3 * (sl & 3);
      bx += q[0]; by += q[1]; bz += q[2];
    }
    // pairs
    const double si = d.topo.sigma[tid];
    const unsigned long long mask_i = d.topo.att_mask[d.topo.type[tid]];
    PT lx = 0, ly = 0, lz = 0, ex = 0, ey = 0, ez = 0;
    lj_gather<PT>(d, P, gi, min(cnt_lj, d.lj_cap), pi, si, mask_i, lx, ly, lz);
    if (gq >= 0) yk_gather<PT>(d, P, gq, qi, min(cnt_q, d.q_cap), pi, ex, ey, ez);
    // src/forces.cpp:141-144: sforce + bforce + (eForce + ljForce)
    const double elx = (double)ex + (double)lx, ely = (double)ey + (double)ly, elz = (double)ez + (double)lz;
    double Fx = sx + bx + elx, Fy = sy + by + ely, Fz = sz + bz + elz;
    const double m = d.topo.mass[tid];
    double vx = d.vx[gi], vy = d.vy[gi], vz = d.vz[gi];
    if (MODE == MODE_FORCES) {
      if (d.out_f) { d.out_f[3 * (size_t)gi] = Fx; d.out_f[3 * (size_t)gi + 1] = Fy; d.out_f[3 * (size_t)gi + 2] = Fz; }
      if (d.out_s) { d.out_s[3 * (size_t)gi] = sx; d.out_s[3 * (size_t)gi + 1] = sy; d.out_s[3 * (size_t)gi + 2] = sz; }
      if (d.out_b) { d.out_b[3 * (size_t)gi] = bx; d.out_b[3 * (size_t)gi + 1] = by; d.out_b[3 * (size_t)gi + 2] = bz; }
      if (d.out_e) { d.out_e[3 * (size_t)gi] = ex; d.out_e[3 * (size_t)gi + 1] = ey; d.out_e[3 * (size_t)gi + 2] = ez; }
      if (d.out_lj) { d.out_lj[3 * (size_t)gi] = lx; d.out_lj[3 * (size_t)gi + 1] = ly; d.out_lj[3 * (size_t)gi + 2] = lz; }
      if (init) { d.fx[gi] = Fx; d.fy[gi] = Fy; d.fz[gi] = Fz; }
    } else if (MODE == MODE_MD) {
      // second half kick (src/md.cpp:451-455)
      const double e = S.expfac, c = S.kick_c / m;
      vx = vx * e + Fx * c; vy = vy * e + Fy * c; vz = vz * e + Fz * c;
      d.vx[gi] = vx; d.vy[gi] = vy; d.vz[gi] = vz;
      d.fx[gi] = Fx; d.fy[gi] = Fy; d.fz[gi] = Fz;
    } else {
      // BD: tforce = F + (fran - mean) + fdrag; v += tforce dt/(2m) (src/md.cpp:462-498)
      const int bl = gi - P.b0;
      double ux, uy, uz;
      bd_draws(S.rng, bl, ux, uy, uz);
      const double amp = P.bd_amp[tid];
      const double rx = amp * ux - S.mean_fran[0], ry = amp * uy - S.mean_fran[1], rz = amp * uz - S.mean_fran[2];
      Fx = Fx + rx + d.gx[gi]; Fy = Fy + ry + d.gy[gi]; Fz = Fz + rz + d.gz[gi];
      const double c = 0.5 * P.dt / m;
      vx = vx + Fx * c; vy = vy + Fy * c; vz = vz + Fz * c;
      d.vx[gi] = vx; d.vy[gi] = vy; d.vz[gi] = vz;
      d.fx[gi] = Fx; d.fy[gi] = Fy; d.fz[gi] = Fz;
    }
    part[0] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    part[1] = vx; part[2] = vy; part[3] = vz;
    if (P.brownian) {
      // noise of the next step (or of the first step at a start) for its mean
      const uint32_t nxt = init ? P.seed : lcg_mulmod(S.rng, P.mult6n);
      double ux, uy, uz;
      bd_draws(nxt, gi - P.b0, ux, uy, uz);
      const double amp = P.bd_amp[tid];
      part[4] = amp * ux; part[5] = amp * uy; part[6] = amp * uz;
    }
    if (rebuild) { part[7] = (double)cnt_lj; part[8] = (double)cnt_q; }
  }
  // per-subunit sums, fixed shape: wave shuffles then the two waves in order
#pragma unroll
  for (int k = 0; k < kPartStride; k++) {
    const double w = wave_sum(part[k]);
    if ((tid & 63) == 0) red[tid >> 6][k] = w;
  }
  __syncthreads();
  if (tid < kPartStride) d.part[(size_t)s * kPartStride + tid] = red[0][tid] + red[1][tid];
}

// One workgroup per replica: fixed-order sums of the per-subunit partials, then the scalar
// bookkeeping of src/md.cpp:400-406, 456-460, 469-478, 506-514.
template <int MODE>
__global__ __launch_bounds__(256) void k_finalize(Dev d, int init, int force_rebuild) {
  __shared__ double red[kPartStride][256];
  const int r = blockIdx.x, tid = threadIdx.x;
  const RepParams& P = d.rp[r];
  RepState& S = d.rs[r];
  for (int k = 0; k < kPartStride; k++) {
    double acc = 0.0;
    for (int s = P.s0 + tid; s < P.s0 + P.ns; s += 256) acc += d.part[(size_t)s * kPartStride + k];
    red[k][tid] = acc;
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
      for (int k = 0; k < kPartStride; k++) red[k][tid] += red[k][tid + o];
    __syncthreads();
  }
  if (tid != 0) return;
  const double ke = red[0][0];
  const int q = P.q_chain;
  const double dt = P.dt;
  if (!P.brownian) {
    if (init) {
      for (int j = 0; j < kMaxChain; j++) { S.xi[j] = 0; S.eta[j] = 0; }
    } else {
      // end of this step: eta update, forward chain with the new kinetic energy
      for (int j = 0; j < q; j++) { S.xi[j] = S.xi_n[j]; S.eta[j] = S.eta_n[j]; }
      for (int j = 0; j < q; j++) if (P.Qb[j] != 0) S.eta[j] = S.eta[j] + 0.5 * dt * S.xi[j];
      for (int j = 0; j < q; j++) chain_xi(j, S.xi, P.Qb, P.dof, P.T, dt, ke);
    }
    // first thermostat half of the next step: reverse chain, eta update
    for (int j = 0; j < kMaxChain; j++) { S.xi_n[j] = S.xi[j]; S.eta_n[j] = S.eta[j]; }
    for (int j = q - 1; j >= 0; j--) chain_xi(j, S.xi_n, P.Qb, P.dof, P.T, dt, ke);
    for (int j = 0; j < q; j++) if (P.Qb[j] != 0) S.eta_n[j] = S.eta_n[j] + 0.5 * dt * S.xi_n[j];
    S.expfac = exp(-0.5 * dt * S.xi_n[0]);
    S.kick_c = 0.5 * dt * sqrt(S.expfac);
  }
  S.ke = ke;
  if (init) {
    S.mean_v[0] = S.mean_v[1] = S.mean_v[2] = 0.0;
    S.vcom = 0.0;
  } else {
    S.mean_v[0] = red[1][0] * P.inv_nb; S.mean_v[1] = red[2][0] * P.inv_nb; S.mean_v[2] = red[3][0] * P.inv_nb;
    S.vcom = sqrt(S.mean_v[0] * S.mean_v[0] + S.mean_v[1] * S.mean_v[1] + S.mean_v[2] * S.mean_v[2]);
    S.step += 1;
  }
  if (P.brownian) {
    S.rng = init ? P.seed : lcg_mulmod(S.rng, P.mult6n);
    S.mean_fran[0] = red[4][0] * P.inv_nb; S.mean_fran[1] = red[5][0] * P.inv_nb; S.mean_fran[2] = red[6][0] * P.inv_nb;
  }
  if (force_rebuild || S.rebuild) {
    if (!init) S.rebuilds += 1;
    S.lj_entries = (long long)red[7][0];
    S.q_entries = (long long)red[8][0];
  }
  S.rebuild = 0;
}

__global__ void k_apply_com(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  const int s = b / d.topo.np;
  const RepState& S = d.rs[d.sub_rep[s]];
  d.vx[b] -= S.mean_v[0]; d.vy[b] -= S.mean_v[1]; d.vz[b] -= S.mean_v[2];
}
__global__ void k_clear_com(Dev d) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.R) return;
  d.rs[r].mean_v[0] = d.rs[r].mean_v[1] = d.rs[r].mean_v[2] = 0.0;
}

// Energies (src/md.cpp:553-593): per bead se, be, ne, ce, ke. Pair energies go to the lower
// bead id of each pair (src/energies.cpp:31-92), using the current lists.
__global__ __launch_bounds__(kBlock) void k_energy(Dev d) {
  extern __shared__ double smem[];
  __shared__ double red[2][5];
  const int np = d.topo.np, nh = d.topo.nh, nq = d.topo.nq;
  double4* own = reinterpret_cast<double4*>(smem);
  double* he = reinterpret_cast<double*>(own + np);
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = d.sub_rep[s];
  const RepParams& P = d.rp[r];
  const double box = P.box, hbox = P.hbox;
  const bool active = tid < np;
  const int gi = s * np + tid;
  if (active) own[tid] = d.pos[gi];
  __syncthreads();
  for (int h = tid; h < nh; h += kBlock) {
    double tmp[12];
    const double c = hinge_forces(own, d.topo.hinge + 4 * h, box, hbox, P.kb, tmp);
    he[h] = P.kb * 0.25 * (1 - fabs(c));  // src/edge.cpp:64-66 after the flip: n0.n1 = |cos|
  }
  __syncthreads();
  double e[5] = {0, 0, 0, 0, 0};
  if (active) {
    const double4 pi = own[tid];
    for (int k = d.topo.e_off[tid]; k < d.topo.e_off[tid + 1]; k++) {  // src/bead.cpp:9-14
      const double4 po = own[d.topo.e_other[k]];
      const double dx = min_img(pi.x - po.x, box, hbox), dy = min_img(pi.y - po.y, box, hbox),
                   dz = min_img(pi.z - po.z, box, hbox);
      const double L = sqrt(dx * dx + dy * dy + dz * dz);
      const double dl = d.topo.e_len0[k] - L;
      e[0] += 0.25 * P.ks * dl * dl;
    }
    for (int k = d.topo.h_off[tid]; k < d.topo.h_off[tid + 1]; k++) e[1] += he[d.topo.h_slot[k] >> 2];
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
        if (rr < P.ecut_el) {
          const double qq = P.qq[qi * nq + qj];
          e[3] += P.yk[qi * nq + qj] * ((qq * exp(-P.kappa * rr)) / rr - (qq * exp(-P.kappa * P.ecut_el)) / P.ecut_el);
        }
      }
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

__global__ void k_bd_uniforms(Dev d, uint32_t state, int nb, double* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double ux, uy, uz;
  bd_draws(state, b, ux, uy, uz);
  out[3 * (size_t)b] = ux; out[3 * (size_t)b + 1] = uy; out[3 * (size_t)b + 2] = uz;
}

// ------------------------------------------------------------------------------------------
static size_t force_smem(const Dev& d) {
  return (size_t)(2 * d.topo.np) * sizeof(double4) + (size_t)d.topo.nh * 12 * sizeof(double);
}

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

int launch_force(const Dev& d, int mode, int init, int force_rebuild, int fp32, hipStream_t st) {
  dim3 g(d.Ns), b(kBlock);
  size_t sm = force_smem(d);
  if (fp32) {
    if (mode == MODE_FORCES) hipLaunchKernelGGL((k_force<MODE_FORCES, float>), g, b, sm, st, d, init, force_rebuild);
    else if (mode == MODE_MD) hipLaunchKernelGGL((k_force<MODE_MD, float>), g, b, sm, st, d, init, force_rebuild);
    else hipLaunchKernelGGL((k_force<MODE_BD, float>), g, b, sm, st, d, init, force_rebuild);
  } else {
    if (mode == MODE_FORCES) hipLaunchKernelGGL((k_force<MODE_FORCES, double>), g, b, sm, st, d, init, force_rebuild);
    else if (mode == MODE_MD) hipLaunchKernelGGL((k_force<MODE_MD, double>), g, b, sm, st, d, init, force_rebuild);
    else hipLaunchKernelGGL((k_force<MODE_BD, double>), g, b, sm, st, d, init, force_rebuild);
  }
  return check_launch();
}

int launch_finalize(const Dev& d, int mode, int init, hipStream_t st) {
  // `init` doubles as the forced-rebuild flag: a start always builds the lists
  hipLaunchKernelGGL(k_finalize<0>, dim3(d.R), dim3(256), 0, st, d, init, init);
  (void)mode;
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

int launch_bd_uniforms_state(uint32_t state, int nb, double* out_dev, const Dev& d, hipStream_t st) {
  hipLaunchKernelGGL(k_bd_uniforms, dim3((nb + 255) / 256), dim3(256), 0, st, d, state, nb, out_dev);
  return check_launch();
}

}  // namespace sf
