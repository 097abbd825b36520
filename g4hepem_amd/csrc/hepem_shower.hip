// hepem_shower.hip -- showers in a 1D stack of slabs (include/g4hepem_amd_shower.h).
//
// The track stack is a pair of structure-of-arrays buffers of `capacity`
// slots. One stepping round over the n tracks in flight:
//   k_sh_safety    isotropic safety = distance to the nearer plane
//   hepem_howfar   (library C-ABI)
//   k_sh_geometry  cut the geometrical step at the next plane along x; the
//                  post-step boundary flag; end point of the straight move
//   hepem_perform  (library C-ABI)
//   k_sh_post      MSC displacement, slab crossing, scoring, leakage, and the
//                  number of stack slots each track needs next round
//   scan           exclusive prefix sum of those counts (3 kernels)
//   k_sh_scatter   survivors, then their secondaries, into the other buffer in
//                  slot order (deterministic), planar with the new count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "g4hepem_amd_shower.h"

namespace {

constexpr int kB = 256;
constexpr double kFix = 1.0e9;           // fixed-point units per MeV
constexpr double kGeomMinLength = 5.0e-8;
constexpr double kElectronMass = 0.51099895000;
constexpr int kMaxRounds = 1 << 20;       // stepping rounds per chunk

#define SH_OK(call)                                   \
  do {                                                \
    if ((call) != hipSuccess) return HEPEM_ERR_HIP;   \
  } while (0)

// splitmix64 finaliser: the stream selector of a primary and of every secondary
__host__ __device__ inline uint64_t Mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t PrimaryStream(int event) { return Mix64(0x5EED000000000000ull ^ (uint64_t)event); }
__host__ __device__ inline uint64_t ChildStream(uint64_t parent, int nstep, int slot) {
  return Mix64(parent ^ Mix64(((uint64_t)(uint32_t)nstep << 2) | (uint64_t)(slot + 1)));
}

// slab geometry, passed by value
struct Slabs {
  int nbins;          // nlayers * nabs
  int nabs;
  int imc[HEPEM_SLAB_MAX_ABS];
  const double* plane;  // nbins + 1 plane positions [mm], plane[0] = 0
};

// one stack buffer
struct Stack {
  hepem_tracks t;     // t.n = tracks in it (stride of the planar fields)
  double* x;
  int32_t* layer;     // slab index
  int32_t* event;     // event index within the chunk
  int32_t* nstep;     // steps taken by the track
};

__global__ void k_sh_fill64(int64_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// primaries of events [e0, e0 + ne): at x = 0, moving along +x, entering slab 0
__global__ void k_sh_init(Stack s, Slabs g, int ptype, double ekin, int e0, int ne) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const hepem_tracks& t = s.t;
  const uint64_t sid = PrimaryStream(e0 + i);
  t.ptype[i] = (int8_t)ptype;
  t.id[i] = (int32_t)(uint32_t)sid;
  t.rng_stream[i] = (uint32_t)(sid >> 32);
  t.imc[i] = g.imc[0];
  t.ekin[i] = ekin;
  t.dir[i] = 1.0;
  t.dir[ne + i] = 0.0;
  t.dir[2 * ne + i] = 0.0;
  t.safety[i] = 0.0;
  t.on_boundary[i] = 1;
  for (int k = 0; k < 4; ++k) t.num_ia_left[k * ne + i] = -1.0;
  t.msc_init_range[i] = 1.0e+21;
  t.msc_dyn_range_factor[i] = 0.04;
  t.msc_tlimit_min[i] = 1.0E-7;
  t.msc_first_step[i] = 1;
  t.rng_ctr[i] = 0u;
  t.rng_gauss[i] = 0.0;
  t.rng_has_gauss[i] = 0;
  s.x[i] = 0.0;
  s.layer[i] = 0;
  s.event[i] = i;
  s.nstep[i] = 0;
}

__global__ void k_sh_safety(Stack s, Slabs g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.t.n) return;
  const int l = s.layer[i];
  const double x = s.x[i];
  s.t.safety[i] = fmax(0.0, fmin(x - g.plane[l], g.plane[l + 1] - x));
}

// Cuts the proposed geometrical step at the next plane along the pre-step
// direction; xend = end point of the straight move, side = +1/-1 when the
// step ends on the upper/lower plane of the slab, 0 inside.
__global__ void k_sh_geometry(Stack s, Slabs g, hepem_howfar_state hs, double* xend, int8_t* side) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = s.t.n;
  if (i >= n) return;
  const int l = s.layer[i];
  const double x = s.x[i], ux = s.t.dir[i];
  const double gs = hs.geom_step[i];
  double dist = 1.0e300;
  int sd = 0;
  if (ux > 0.0) {
    dist = (g.plane[l + 1] - x) / ux;
    sd = 1;
  } else if (ux < 0.0) {
    dist = (g.plane[l] - x) / ux;
    sd = -1;
  }
  if (hs.winner[i] != -99 && dist < gs) {
    hs.geom_step[i] = dist;
    s.t.on_boundary[i] = 1;
    side[i] = (int8_t)sd;
    xend[i] = sd > 0 ? g.plane[l + 1] : g.plane[l];
  } else {
    s.t.on_boundary[i] = 0;
    side[i] = 0;
    xend[i] = fmin(fmax(x + gs * ux, g.plane[l]), g.plane[l + 1]);
  }
}

// After Perform: position, slab crossing, scoring and next-round slot count.
// keep[i] = 1 if the track itself goes on, nsk[i] = secondaries to stack;
// cnt[i] = keep[i] + nsk[i].
__global__ void k_sh_post(Stack s, Slabs g, hepem_step_out res, const double* __restrict__ xend,
                          const int8_t* __restrict__ side, unsigned long long* dep, unsigned long long* leak,
                          uint8_t* keep, uint8_t* nsk, uint32_t* cnt, unsigned long long* nsecTotal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = s.t.n;
  if (i >= n) return;
  const hepem_tracks& t = s.t;
  const int ev = s.event[i];
  int l = s.layer[i];
  const double edep = res.edep[i];
  if (edep > 0.0) atomicAdd(&dep[(size_t)ev * g.nbins + l], (unsigned long long)llrint(edep * kFix));
  double x = xend[i];
  const int sd = side[i];
  const bool rejected = res.winner[i] == -99;
  const bool isE = t.ptype[i] != 2;
  if (sd == 0 && isE && !rejected) {
    // MSC displacement (G4HepEmTrackingManager.cc:527-566) with the slab safety
    const double d0 = res.displacement[i], d1 = res.displacement[n + i], d2 = res.displacement[2 * n + i];
    const double dLength2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (dLength2 > kGeomMinLength * kGeomMinLength) {
      const double dispR = sqrt(dLength2);
      const double postSafety = 0.99 * fmax(0.0, fmin(x - g.plane[l], g.plane[l + 1] - x));
      if (postSafety > 0.0 && dispR <= postSafety) {
        x += d0;
      } else if (dispR < postSafety) {
        x += d0;
      } else if (postSafety > kGeomMinLength) {
        x += d0 * (postSafety / dispR);
      }
    }
  }
  const double ekin = t.ekin[i];
  bool alive = !rejected && ekin > 0.0;
  bool left = false;
  if (alive && sd != 0) {
    l += sd;
    if (l < 0 || l >= g.nbins) {
      // leaves the stack: its kinetic energy (and an e+'s annihilation energy) is lost
      const double e = ekin + (t.ptype[i] == 1 ? 2.0 * kElectronMass : 0.0);
      atomicAdd(&leak[ev], (unsigned long long)llrint(e * kFix));
      alive = false;
      left = true;
      l -= sd;  // keep the index in range for the record
    }
  }
  s.x[i] = x;
  s.layer[i] = l;
  // a track that crossed a plane had no interaction, so a track leaving the
  // stack has no secondaries; the guard keeps that an invariant of the stack
  const int ns = (rejected || left) ? 0 : res.nsec[i];
  keep[i] = alive ? 1 : 0;
  nsk[i] = (uint8_t)ns;
  cnt[i] = (alive ? 1u : 0u) + (uint32_t)ns;
  if (ns > 0) atomicAdd(nsecTotal, (unsigned long long)ns);
}

// exclusive scan of cnt[0, n) into offs; bsum[b] = sum of block b
__global__ void k_scan_block(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ offs, uint32_t* __restrict__ bsum, int n) {
  __shared__ uint32_t sh[kB];
  const int i = blockIdx.x * kB + threadIdx.x;
  const uint32_t v = i < n ? cnt[i] : 0u;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < kB; o <<= 1) {
    const uint32_t a = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += a;
    __syncthreads();
  }
  if (i < n) offs[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == kB - 1) bsum[blockIdx.x] = sh[kB - 1];
}
// in-place exclusive scan of the nb block sums by one block; total -> *total
__global__ void k_scan_sums(uint32_t* bsum, int nb, uint32_t* total) {
  __shared__ uint32_t sh[kB];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0u;
  __syncthreads();
  for (int base = 0; base < nb; base += kB) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0u;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kB; o <<= 1) {
      const uint32_t a = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
      __syncthreads();
      sh[threadIdx.x] += a;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == kB - 1) carry += sh[kB - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}
__global__ void k_scan_add(uint32_t* offs, const uint32_t* __restrict__ bsum, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) offs[i] += bsum[blockIdx.x];
}

// Survivor i goes to slot offs[i], its secondaries to the following slots.
// Writes nothing beyond `cap` (the host then reports the overflow).
__global__ void k_sh_scatter(Stack a, Stack b, Slabs g, hepem_step_out res, const uint8_t* __restrict__ keep,
                             const uint8_t* __restrict__ nsk, const uint32_t* __restrict__ offs,
                             const uint32_t* __restrict__ total, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.t.n;
  if (i >= n) return;
  const int m = (int)*total;
  if (m > cap) return;
  const hepem_tracks& s = a.t;
  const hepem_tracks& d = b.t;
  int j = (int)offs[i];
  const uint64_t sid = ((uint64_t)s.rng_stream[i] << 32) | (uint32_t)s.id[i];
  const int ns1 = a.nstep[i] + 1;
  if (keep[i]) {
    d.ptype[j] = s.ptype[i];
    d.id[j] = s.id[i];
    d.rng_stream[j] = s.rng_stream[i];
    d.imc[j] = g.imc[a.layer[i] % g.nabs];
    d.ekin[j] = s.ekin[i];
    for (int c = 0; c < 3; ++c) d.dir[c * m + j] = s.dir[c * n + i];
    d.safety[j] = 0.0;
    d.on_boundary[j] = s.on_boundary[i];
    for (int k = 0; k < 4; ++k) d.num_ia_left[k * m + j] = s.num_ia_left[k * n + i];
    d.msc_init_range[j] = s.msc_init_range[i];
    d.msc_dyn_range_factor[j] = s.msc_dyn_range_factor[i];
    d.msc_tlimit_min[j] = s.msc_tlimit_min[i];
    d.msc_first_step[j] = s.msc_first_step[i];
    d.rng_ctr[j] = s.rng_ctr[i];
    d.rng_gauss[j] = s.rng_gauss[i];
    d.rng_has_gauss[j] = s.rng_has_gauss[i];
    b.x[j] = a.x[i];
    b.layer[j] = a.layer[i];
    b.event[j] = a.event[i];
    b.nstep[j] = ns1;
    ++j;
  }
  const int nsec = nsk[i];
  for (int sl = 0; sl < 2; ++sl) {
    if (sl >= nsec) break;
    const uint64_t cid = ChildStream(sid, ns1, sl);
    d.ptype[j] = res.sec_ptype[sl * n + i];
    d.id[j] = (int32_t)(uint32_t)cid;
    d.rng_stream[j] = (uint32_t)(cid >> 32);
    d.imc[j] = g.imc[a.layer[i] % g.nabs];
    d.ekin[j] = res.sec_ekin[sl * n + i];
    for (int c = 0; c < 3; ++c) d.dir[c * m + j] = res.sec_dir[(c * 2 + sl) * n + i];
    d.safety[j] = 0.0;
    d.on_boundary[j] = 0;
    for (int k = 0; k < 4; ++k) d.num_ia_left[k * m + j] = -1.0;
    d.msc_init_range[j] = 1.0e+21;
    d.msc_dyn_range_factor[j] = 0.04;
    d.msc_tlimit_min[j] = 1.0E-7;
    d.msc_first_step[j] = 1;
    d.rng_ctr[j] = 0u;
    d.rng_gauss[j] = 0.0;
    d.rng_has_gauss[j] = 0;
    b.x[j] = a.x[i];
    b.layer[j] = a.layer[i];
    b.event[j] = a.event[i];
    b.nstep[j] = 0;
    ++j;
  }
}

// device memory of one run, freed on every exit path
struct Arena {
  std::vector<void*> ptrs;
  ~Arena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T> bool get(T*& p, size_t n) {
    void* v = nullptr;
    if (hipMalloc(&v, (n ? n : 1) * sizeof(T)) != hipSuccess) return false;
    ptrs.push_back(v);
    p = static_cast<T*>(v);
    return true;
  }
};

bool AllocStack(Arena& A, Stack& s, int cap) {
  hepem_tracks& t = s.t;
  t.n = 0;
  return A.get(t.ptype, cap) && A.get(t.id, cap) && A.get(t.imc, cap) && A.get(t.ekin, cap) && A.get(t.dir, 3 * (size_t)cap) &&
         A.get(t.safety, cap) && A.get(t.on_boundary, cap) && A.get(t.num_ia_left, 4 * (size_t)cap) &&
         A.get(t.msc_init_range, cap) && A.get(t.msc_dyn_range_factor, cap) && A.get(t.msc_tlimit_min, cap) &&
         A.get(t.msc_first_step, cap) && A.get(t.rng_ctr, cap) && A.get(t.rng_gauss, cap) && A.get(t.rng_has_gauss, cap) &&
         A.get(t.rng_stream, cap) && A.get(s.x, cap) && A.get(s.layer, cap) && A.get(s.event, cap) && A.get(s.nstep, cap);
}

// Kahan-compensated accumulator
struct Kahan {
  double s = 0.0, c = 0.0;
  void add(double v) {
    const double y = v - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

}  // namespace

extern "C" int hepem_slab_run(const hepem_ctx* ctx, const hepem_slab_setup* su, int32_t nevents, uint64_t seed,
                              hepem_slab_result* r, void* stream) {
  if (!ctx || !su || !r || !r->edep_sum || !r->edep_sum2 || nevents < 0) return HEPEM_ERR_INVALID;
  if (su->nlayers < 1 || su->nabs < 1 || su->nabs > HEPEM_SLAB_MAX_ABS || su->capacity < 1 || su->events_per_chunk < 1 ||
      !(su->ekin0 > 0.0) || su->primary < 0 || su->primary > 2)
    return HEPEM_ERR_INVALID;
  for (int a = 0; a < su->nabs; ++a)
    if (!(su->thickness[a] > 0.0) || su->imc[a] < 0) return HEPEM_ERR_INVALID;
  const int nbins = su->nlayers * su->nabs;
  const int cap = su->capacity;
  const int chunk = su->events_per_chunk < cap ? su->events_per_chunk : cap;
  const hipStream_t st = static_cast<hipStream_t>(stream);

  std::vector<double> planes((size_t)nbins + 1);
  planes[0] = 0.0;
  for (int k = 0; k < nbins; ++k) planes[k + 1] = planes[k] + su->thickness[k % su->nabs];

  Arena A;
  Stack s[2];
  hepem_howfar_state hs;
  hepem_step_out res;
  double *dPlanes, *xend;
  int8_t* side;
  uint8_t *keep, *nsk, *work;
  uint32_t *cnt, *offs, *bsum, *total;
  unsigned long long *dep, *leak, *nsec;
  const int nblk = (cap + kB - 1) / kB;
  if (!AllocStack(A, s[0], cap) || !AllocStack(A, s[1], cap) || !A.get(hs.geom_step, cap) || !A.get(hs.range, cap) ||
      !A.get(hs.mfp, 4 * (size_t)cap) || !A.get(hs.winner, cap) || !A.get(hs.msc_lambtr1, cap) ||
      !A.get(hs.msc_true_step, cap) || !A.get(hs.msc_zpath, cap) || !A.get(hs.msc_par1, cap) || !A.get(hs.msc_par2, cap) ||
      !A.get(hs.msc_par3, cap) || !A.get(hs.msc_active, cap) || !A.get(hs.msc_displace, cap) || !A.get(hs.pe_xsec, cap) ||
      !A.get(res.true_step, cap) || !A.get(res.geom_step, cap) || !A.get(res.edep, cap) ||
      !A.get(res.displacement, 3 * (size_t)cap) || !A.get(res.winner, cap) || !A.get(res.stopped, cap) ||
      !A.get(res.nsec, cap) || !A.get(res.sec_ptype, 2 * (size_t)cap) || !A.get(res.sec_ekin, 2 * (size_t)cap) ||
      !A.get(res.sec_dir, 6 * (size_t)cap) || !A.get(dPlanes, planes.size()) || !A.get(xend, cap) || !A.get(side, cap) ||
      !A.get(keep, cap) || !A.get(nsk, cap) || !A.get(cnt, cap) || !A.get(offs, cap) || !A.get(bsum, nblk) || !A.get(total, 1) ||
      !A.get(dep, (size_t)chunk * nbins) || !A.get(leak, chunk) || !A.get(nsec, 1) ||
      !A.get(work, hepem_step_work_bytes(cap)))
    return HEPEM_ERR_ALLOC;
  SH_OK(hipMemcpy(dPlanes, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice));

  Slabs g;
  g.nbins = nbins;
  g.nabs = su->nabs;
  for (int a = 0; a < HEPEM_SLAB_MAX_ABS; ++a) g.imc[a] = a < su->nabs ? su->imc[a] : 0;
  g.plane = dPlanes;

  std::vector<Kahan> ks(nbins), ks2(nbins);
  Kahan kleak, ktot;
  std::vector<unsigned long long> hdep((size_t)chunk * nbins), hleak(chunk);
  r->steps = 0;
  r->tracks = 0;
  r->iterations = 0;
  r->max_in_flight = 0;
  unsigned long long hnsec = 0;
  hipLaunchKernelGGL(k_sh_fill64, dim3(1), dim3(kB), 0, st, reinterpret_cast<int64_t*>(nsec), 1);

  for (int e0 = 0; e0 < nevents; e0 += chunk) {
    const int ne = (nevents - e0) < chunk ? (nevents - e0) : chunk;
    hipLaunchKernelGGL(k_sh_fill64, dim3((ne * nbins + kB - 1) / kB), dim3(kB), 0, st, reinterpret_cast<int64_t*>(dep),
                       ne * nbins);
    hipLaunchKernelGGL(k_sh_fill64, dim3((ne + kB - 1) / kB), dim3(kB), 0, st, reinterpret_cast<int64_t*>(leak), ne);
    int cur = 0;
    s[cur].t.n = ne;
    hipLaunchKernelGGL(k_sh_init, dim3((ne + kB - 1) / kB), dim3(kB), 0, st, s[cur], g, su->primary, su->ekin0, e0, ne);
    SH_OK(hipGetLastError());
    r->tracks += ne;
    int n = ne;
    int rounds = 0;
    while (n > 0) {
      // every step of an e-/e+ loses energy and every gamma step ends on a plane
      // or in an interaction, so a chunk ends; the bound only turns a broken
      // invariant into an error instead of an endless loop
      if (++rounds > kMaxRounds) return HEPEM_ERR_CAPACITY;
      if (n > r->max_in_flight) r->max_in_flight = n;
      Stack& a = s[cur];
      Stack& b = s[cur ^ 1];
      a.t.n = n;
      hs.n = n;
      const int grid = (n + kB - 1) / kB;
      hipLaunchKernelGGL(k_sh_safety, dim3(grid), dim3(kB), 0, st, a, g);
      int rc = hepem_howfar(ctx, seed, &a.t, &hs, work, stream);
      if (rc != HEPEM_OK) return rc;
      hipLaunchKernelGGL(k_sh_geometry, dim3(grid), dim3(kB), 0, st, a, g, hs, xend, side);
      rc = hepem_perform(ctx, seed, &a.t, &hs, &res, work, stream);
      if (rc != HEPEM_OK) return rc;
      hipLaunchKernelGGL(k_sh_post, dim3(grid), dim3(kB), 0, st, a, g, res, xend, side, dep, leak, keep, nsk, cnt, nsec);
      hipLaunchKernelGGL(k_scan_block, dim3(grid), dim3(kB), 0, st, cnt, offs, bsum, n);
      hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kB), 0, st, bsum, grid, total);
      hipLaunchKernelGGL(k_scan_add, dim3(grid), dim3(kB), 0, st, offs, bsum, n);
      hipLaunchKernelGGL(k_sh_scatter, dim3(grid), dim3(kB), 0, st, a, b, g, res, keep, nsk, offs, total, cap);
      SH_OK(hipGetLastError());
      uint32_t m = 0;
      SH_OK(hipMemcpyAsync(&m, total, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      SH_OK(hipStreamSynchronize(st));
      r->steps += n;
      r->iterations += 1;
      if (m > (uint32_t)cap) return HEPEM_ERR_CAPACITY;
      n = (int)m;
      cur ^= 1;
    }
    SH_OK(hipMemcpyAsync(hdep.data(), dep, (size_t)ne * nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    SH_OK(hipMemcpyAsync(hleak.data(), leak, (size_t)ne * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    SH_OK(hipStreamSynchronize(st));
    if (r->edep_event)
      for (size_t k = 0; k < (size_t)ne * nbins; ++k) r->edep_event[(size_t)e0 * nbins + k] = (int64_t)hdep[k];
    if (r->leak_event)
      for (int e = 0; e < ne; ++e) r->leak_event[e0 + e] = (int64_t)hleak[e];
    for (int e = 0; e < ne; ++e) {
      for (int k = 0; k < nbins; ++k) {
        const double v = (double)hdep[(size_t)e * nbins + k] / kFix;
        ks[k].add(v);
        ks2[k].add(v * v);
        ktot.add(v);
      }
      kleak.add((double)hleak[e] / kFix);
    }
  }
  SH_OK(hipMemcpyAsync(&hnsec, nsec, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  SH_OK(hipStreamSynchronize(st));
  r->tracks += (int64_t)hnsec;
  for (int k = 0; k < nbins; ++k) {
    r->edep_sum[k] = ks[k].s;
    r->edep_sum2[k] = ks2[k].s;
  }
  r->leak_sum = kleak.s;
  r->edep_total = ktot.s;
  return HEPEM_OK;
}
