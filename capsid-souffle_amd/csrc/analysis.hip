// Oligomers (clusters) of subunits on the GPU: two subunits of one replica are joined when any
// of their beads are closer than a contact distance under the single-shift minimum image, and
// an oligomer is a connected component of that relation. This is the mass-spectrum analysis of
// src/analysis.cpp:143-186 (contact 2 sigma there), which grows each oligomer breadth-first
// from its smallest unassigned subunit; the components do not depend on how they are found, so
// here every workgroup tests its subunit against the higher-numbered subunits of its replica
// and joins contacting pairs in a lock-free union-find whose roots are the smallest members.
//   k_olig_extent    largest bead distance from its centre bead, per replica (one thread per bead)
//   k_olig_contacts  one 128-thread workgroup per subunit: centre-distance candidates, then all
//                    bead pairs of each candidate against the subunit's own beads
//   k_olig_label     root of each subunit (replica-local label) and oligomer sizes
//   k_olig_spectrum  oligomers counted by size, per replica
// Results are exact and order-independent: labels are the smallest subunit index of each
// oligomer whatever order the unions ran in.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "sf_internal.hpp"

namespace sf {

namespace {

__device__ __forceinline__ double img(double v, double box, double hbox) {
  // single conditional shift per axis, as src/functions.h:42-56
  if (v > hbox) v -= box;
  else if (v < -hbox) v += box;
  return v;
}

__device__ __forceinline__ int load_parent(const int* p, int x) {
  return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Parent pointers only ever point to smaller indices, so the walk ends at the root.
__device__ int find_root(const int* p, int x) {
  int y = load_parent(p, x);
  while (y != x) {
    x = y;
    y = load_parent(p, x);
  }
  return x;
}

// Hooks the larger of the two roots under the smaller one; a failed exchange means that root
// was hooked meanwhile, and the walk starts again from a strictly smaller root.
__device__ void unite(int* p, int a, int b) {
  while (true) {
    a = find_root(p, a);
    b = find_root(p, b);
    if (a == b) return;
    if (a > b) { const int t = a; a = b; b = t; }
    if (atomicCAS(p + b, b, a) == b) return;
  }
}

__global__ void k_olig_extent(Dev d, unsigned long long* ext) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.Nb) return;
  const int s = b / d.topo.np;
  const RepParams& P = d.rp[d.sub_rep[s]];
  const double4 p = d.pos[b], c = d.centres[s];
  const double dx = img(p.x - c.x, P.box, P.hbox), dy = img(p.y - c.y, P.box, P.hbox), dz = img(p.z - c.z, P.box, P.hbox);
  const double r2 = dx * dx + dy * dy + dz * dz;
  // non-negative doubles order like their bit patterns
  atomicMax(ext + d.sub_rep[s], (unsigned long long)__double_as_longlong(r2));
}

// Resets the union-find and the counts; ext (one slot per replica, R <= ns) and the total are
// cleared here too, so the analysis issues no memset.
__global__ void k_olig_init(int ns, int R, int* parent, int* size, int* spectrum, unsigned long long* ext, int* total) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  parent[s] = s;
  size[s] = 0;
  spectrum[s] = 0;
  if (s < R) ext[s] = 0ull;
  if (s == 0) *total = 0;
}

__global__ __launch_bounds__(kBlock) void k_olig_contacts(Dev d, double contact, const unsigned long long* ext,
                                                           int* parent) {
  __shared__ double4 other[kBlock];
  __shared__ int cand[kBlock];
  __shared__ int ncand, hit;
  const int s = blockIdx.x, tid = threadIdx.x;
  const int r = d.sub_rep[s];
  const RepParams& P = d.rp[r];
  const int np = d.topo.np, end = P.s0 + P.ns;
  if (s + 1 >= end) return;  // uniform: no higher-numbered subunit in the replica
  const double box = P.box, hbox = P.hbox;
  const double c2 = contact * contact;
  // centres closer than contact + 2 x (largest bead distance from a centre bead) may touch
  const double e = sqrt(__longlong_as_double((long long)ext[r]));
  const double reach = contact + 2.0 * e + 1e-9;
  const double reach2 = reach * reach;
  const bool active = tid < np;
  double4 own = make_double4(0, 0, 0, 0);
  if (active) own = d.pos[(size_t)s * np + tid];
  const double4 cs = d.centres[s];
  for (int base = s + 1; base < end; base += kBlock) {
    if (tid == 0) ncand = 0;
    __syncthreads();
    const int t = base + tid;
    if (t < end) {
      const double4 ct = d.centres[t];
      const double dx = img(cs.x - ct.x, box, hbox), dy = img(cs.y - ct.y, box, hbox), dz = img(cs.z - ct.z, box, hbox);
      if (dx * dx + dy * dy + dz * dz < reach2) cand[atomicAdd(&ncand, 1)] = t;
    }
    __syncthreads();
    const int n = ncand;
    for (int k = 0; k < n; k++) {
      const int u = cand[k];
      if (active) other[tid] = d.pos[(size_t)u * np + tid];
      if (tid == 0) hit = 0;
      __syncthreads();
      if (active) {
        for (int m = 0; m < np; m++) {
          const double4 q = other[m];
          const double dx = img(own.x - q.x, box, hbox), dy = img(own.y - q.y, box, hbox), dz = img(own.z - q.z, box, hbox);
          if (dx * dx + dy * dy + dz * dz < c2) { hit = 1; break; }
        }
      }
      __syncthreads();
      if (tid == 0 && hit) unite(parent, s, u);
      __syncthreads();  // `other` and `hit` are reused by the next candidate
    }
  }
}

__global__ void k_olig_label(Dev d, const int* parent, int* label, int* size) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.Ns) return;
  const int root = find_root(parent, s);
  label[s] = root - d.rp[d.sub_rep[s]].s0;
  atomicAdd(size + root, 1);
}

// spectrum[s0 + k - 1] = oligomers of k subunits in the replica starting at subunit s0
__global__ void k_olig_spectrum(Dev d, const int* label, const int* size, int* spectrum, int* total) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.Ns) return;
  const int s0 = d.rp[d.sub_rep[s]].s0;
  if (label[s] == s - s0) {
    atomicAdd(spectrum + s0 + size[s] - 1, 1);
    atomicAdd(total, 1);
  }
}

}  // namespace

// Fixed device buffers of the allocation-free entry point (sf_oligomers_noalloc): ext takes
// the first kOligoScratchSubunits words (R <= Ns), then parent, size, label and spectrum
// (kOligoScratchSubunits ints each), then the total. Every GPU the library runs on holds its own
// copy until the process ends. In host code `scratch` names a host-side stand-in, not device
// memory: kernels and copies get the device address from hipGetSymbolAddress instead.
constexpr size_t kOligoScratchWords = 3 * (size_t)kOligoScratchSubunits + 32;
static __device__ __align__(256) unsigned long long scratch[kOligoScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each device, once resolved
}  // namespace

int launch_oligomers(const Dev& d, double contact, OligoBuf& o, hipStream_t st) {
  const int nb_blocks = (d.Nb + 255) / 256, ns_blocks = (d.Ns + 255) / 256;
  k_olig_init<<<ns_blocks, 256, 0, st>>>(d.Ns, d.R, o.parent, o.size, o.spectrum, o.ext, o.total);
  k_olig_extent<<<nb_blocks, 256, 0, st>>>(d, o.ext);
  k_olig_contacts<<<d.Ns, kBlock, 0, st>>>(d, contact, o.ext, o.parent);
  k_olig_label<<<ns_blocks, 256, 0, st>>>(d, o.parent, o.label, o.size);
  k_olig_spectrum<<<ns_blocks, 256, 0, st>>>(d, o.label, o.size, o.spectrum, o.total);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("oligomer kernels: ") + hipGetErrorString(e));
    return SF_E_HIP;
  }
  return SF_OK;
}

int oligo_scratch(int dev, int ns, int R, OligoBuf& o) {
  if (ns <= 0 || R <= 0 || R > ns) { set_error("bad handle sizes (oligomers)"); return SF_E_ARG; }
  if (ns > kOligoScratchSubunits) {
    set_error("sf_oligomers_noalloc serves at most " + std::to_string(kOligoScratchSubunits) +
              " subunits per handle; this one has " + std::to_string(ns) + " (use sf_oligomers)");
    return SF_E_ARG;
  }
  if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range (oligomers)"); return SF_E_ARG; }
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      // the address is per device: resolve it with that device current
      int cur = -1;
      if (hipGetDevice(&cur) != hipSuccess) { set_error("hipGetDevice failed (oligomers)"); return SF_E_HIP; }
      if (cur != dev && hipSetDevice(dev) != hipSuccess) { set_error("hipSetDevice failed (oligomers)"); return SF_E_HIP; }
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (cur != dev) (void)hipSetDevice(cur);
      if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        return SF_E_HIP;
      }
      g_scratch_ptr[dev] = ptr;
    }
    base = g_scratch_ptr[dev];
  }
  unsigned long long* w = static_cast<unsigned long long*>(base);
  int* ints = reinterpret_cast<int*>(w + kOligoScratchSubunits);
  o.ext = w;
  o.parent = ints;
  o.size = ints + kOligoScratchSubunits;
  o.label = ints + 2 * (size_t)kOligoScratchSubunits;
  o.spectrum = ints + 3 * (size_t)kOligoScratchSubunits;
  o.total = ints + 4 * (size_t)kOligoScratchSubunits;
  return SF_OK;
}

}  // namespace sf
