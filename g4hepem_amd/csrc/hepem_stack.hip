// hepem_stack.hip -- a device track stack for callers that step batches with
// their own geometry (include/g4hepem_amd.h, hepem_stack_*): the next batch
// from a step's results, as G4HepEmTrackingManager::StackSecondaries stacks
// one track's secondaries (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:
// 1254-1330), for every track of the batch in three kernels:
//   k_st_count    per parent slot: whether the track stays, which secondaries
//                 are stacked and the energy the production cuts deposit; the
//                 block's exclusive scans of the e-/e+ and gamma slots
//   k_st_scan     exclusive scans of the block sums (one block) and the
//                 next batch's counts
//   k_st_scatter  survivors and stacked secondaries into the other buffer,
//                 e-/e+ first, then gammas, each in parent order
// and k_st_append for hepem_stack_push.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <type_traits>
#include <vector>

#include "device/dev_common.hh"
#include "device/dev_scan.hh"
#include "g4hepem_amd.h"
#include "hepem_internal.hh"

namespace {

constexpr int kB = 256;          // threads of the per-slot kernels
constexpr int kScanB = 1024;     // threads of the one-block scan
constexpr double kElectronMass = 0.51099895000;  // CLHEP::electron_mass_c2 [MeV]

// production thresholds of a couple and whether its region applies them
struct CoupleCuts {
  double e, p, g;
  int apply;
};

// splitmix64 finaliser (the stream of a stacked secondary is derived from its
// parent's 64-bit stream, the parent's post-step draw counter and the slot)
__device__ inline uint64_t Mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline uint32_t StreamHi(const hepem_tracks& t, int i) { return t.rng_stream ? t.rng_stream[i] : 0u; }

// one buffer of tracks: every hepem_tracks field for `cap` tracks
bool AllocTracks(hepem_tracks& t, int cap) {
  t = hepem_tracks{};
  const size_t c = (size_t)(cap > 0 ? cap : 1);
  auto get = [&](auto*& p, size_t k) {
    void* v = nullptr;
    if (hipMalloc(&v, k * c * sizeof(*p)) != hipSuccess) return false;
    p = static_cast<std::remove_reference_t<decltype(p)>>(v);
    return true;
  };
  return get(t.ptype, 1) && get(t.id, 1) && get(t.imc, 1) && get(t.ekin, 1) && get(t.dir, 3) && get(t.safety, 1) &&
         get(t.on_boundary, 1) && get(t.num_ia_left, 4) && get(t.msc_init_range, 1) && get(t.msc_dyn_range_factor, 1) &&
         get(t.msc_tlimit_min, 1) && get(t.msc_first_step, 1) && get(t.rng_ctr, 1) && get(t.rng_gauss, 1) &&
         get(t.rng_has_gauss, 1) && get(t.rng_stream, 1);
}

void FreeTracks(hepem_tracks& t) {
  void* ps[] = {t.ptype, t.id, t.imc, t.ekin, t.dir, t.safety, t.on_boundary, t.num_ia_left, t.msc_init_range,
                t.msc_dyn_range_factor, t.msc_tlimit_min, t.msc_first_step, t.rng_ctr, t.rng_gauss, t.rng_has_gauss,
                t.rng_stream};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  t = hepem_tracks{};
}

// track i of s (planar stride s.n) to slot j of d (planar stride d.n)
__device__ inline void CopyTrack(const hepem_tracks& s, int i, const hepem_tracks& d, int j) {
  const int n = s.n, m = d.n;
  d.ptype[j] = s.ptype[i];
  d.id[j] = s.id[i];
  if (d.rng_stream) d.rng_stream[j] = StreamHi(s, i);
  d.imc[j] = s.imc[i];
  d.ekin[j] = s.ekin[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) d.dir[c * m + j] = s.dir[c * n + i];
  d.safety[j] = s.safety[i];
  d.on_boundary[j] = s.on_boundary[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) d.num_ia_left[k * m + j] = s.num_ia_left[k * n + i];
  d.msc_init_range[j] = s.msc_init_range[i];
  d.msc_dyn_range_factor[j] = s.msc_dyn_range_factor[i];
  d.msc_tlimit_min[j] = s.msc_tlimit_min[i];
  d.msc_first_step[j] = s.msc_first_step[i];
  d.rng_ctr[j] = s.rng_ctr[i];
  d.rng_gauss[j] = s.rng_gauss[i];
  d.rng_has_gauss[j] = s.rng_has_gauss[i];
}

// Slots [0, a.n) of the new buffer d from a, [a.n, a.n + b.n) from b.
__global__ void __launch_bounds__(kB) k_st_append(hepem_tracks a, hepem_tracks b, hepem_tracks d) {
  const int j = blockIdx.x * kB + threadIdx.x;
  if (j >= d.n) return;
  if (j < a.n) CopyTrack(a, j, d, j);
  else CopyTrack(b, j - a.n, d, j);
}

// Per parent slot i of the step's output `out` / `res`: alive[i] (the track
// stays), smask[i] (bit s: secondary slot s is stacked), the e-/e+ and gamma
// slots it needs in the next batch and their exclusive scans within the
// block (offE, offG; block sums in bsumE, bsumG), and the energy the cuts
// deposit (cutEdep, optional).
__global__ void __launch_bounds__(kB) k_st_count(hepem_tracks out, hepem_step_out res, const uint8_t* __restrict__ keep,
                                                 const CoupleCuts* __restrict__ cuts, int nmc, uint8_t* __restrict__ alive,
                                                 uint8_t* __restrict__ smask, uint32_t* __restrict__ offE,
                                                 uint32_t* __restrict__ offG, uint32_t* __restrict__ bsumE,
                                                 uint32_t* __restrict__ bsumG, double* __restrict__ cutEdep) {
  const int i = blockIdx.x * kB + threadIdx.x;
  const int n = out.n;
  const bool valid = i < n;
  uint32_t cE = 0u, cG = 0u;
  if (valid) {
    const bool stays = (keep == nullptr || keep[i] != 0) && out.ekin[i] > 0.0;
    const int imc = out.imc[i];
    const bool apply = imc >= 0 && imc < nmc && cuts[imc].apply != 0;
    const int nsec = min(max(res.nsec[i], 0), 2);
    double dep = 0.0;
    uint32_t mask = 0u;
    for (int s = 0; s < nsec; ++s) {
      const int8_t pt = res.sec_ptype[s * n + i];
      const double se = res.sec_ekin[s * n + i];
      bool cut = false;
      if (apply) {
        const CoupleCuts c = cuts[imc];
        if (pt == HEPEM_ELECTRON) cut = se < c.e;
        else if (pt == HEPEM_POSITRON) cut = kElectronMass < c.g && se < c.p;
        else cut = se < c.g;
      }
      if (cut) {
        dep += pt == HEPEM_POSITRON ? se + 2.0 * kElectronMass : se;
      } else {
        mask |= 1u << s;
        if (pt == HEPEM_GAMMA) ++cG;
        else ++cE;
      }
    }
    if (stays) {
      if (out.ptype[i] == HEPEM_GAMMA) ++cG;
      else ++cE;
    }
    alive[i] = stays ? 1 : 0;
    smask[i] = (uint8_t)mask;
    if (cutEdep) cutEdep[i] = dep;
  }
  uint32_t xE, xG, tE, tG, tw;
  hepem::BlockScan2<kB>(cE, cG, 0u, xE, xG, tE, tG, tw);
  if (valid) {
    offE[i] = xE;
    offG[i] = xG;
  }
  if (threadIdx.x == 0) {
    bsumE[blockIdx.x] = tE;
    bsumG[blockIdx.x] = tG;
  }
}

// In-place exclusive scans of the nb block sums (one block); total[0] = slots
// of the next batch, total[1] = its e-/e+ part.
__global__ void __launch_bounds__(kScanB) k_st_scan(uint32_t* bsumE, uint32_t* bsumG, int nb, uint32_t* total) {
  __shared__ uint32_t carryE, carryG;
  if (threadIdx.x == 0) {
    carryE = 0u;
    carryG = 0u;
  }
  __syncthreads();
  for (int base = 0; base < nb; base += kScanB) {
    const int i = base + threadIdx.x;
    const uint32_t a = i < nb ? bsumE[i] : 0u;
    const uint32_t b = i < nb ? bsumG[i] : 0u;
    uint32_t xa, xb, ta, tb, tw;
    hepem::BlockScan2<kScanB>(a, b, 0u, xa, xb, ta, tb, tw);
    if (i < nb) {
      bsumE[i] = carryE + xa;
      bsumG[i] = carryG + xb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carryE += ta;
      carryG += tb;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    total[0] = carryE + carryG;
    total[1] = carryE;
  }
}

// Survivor i and its stacked secondaries into d: e-/e+ from bsumE[block] +
// offE[i], gammas from total[1] + bsumG[block] + offG[i], in the order
// survivor, secondary 0, secondary 1. Writes nothing when the next batch
// exceeds cap (d.n is the host's bound; the host reads the count and reports
// it).
__global__ void __launch_bounds__(kB) k_st_scatter(hepem_tracks out, hepem_step_out res, const uint8_t* __restrict__ alive,
                                                   const uint8_t* __restrict__ smask, const uint32_t* __restrict__ offE,
                                                   const uint32_t* __restrict__ offG, const uint32_t* __restrict__ bsumE,
                                                   const uint32_t* __restrict__ bsumG, const uint32_t* __restrict__ total,
                                                   hepem_tracks d, int cap, int32_t* __restrict__ parent,
                                                   int8_t* __restrict__ origin) {
  const int i = blockIdx.x * kB + threadIdx.x;
  const int n = out.n;
  if (i >= n) return;
  const int m = (int)total[0];
  if (m > cap) return;
  d.n = m;
  int jE = (int)(bsumE[blockIdx.x] + offE[i]);
  int jG = (int)(total[1] + bsumG[blockIdx.x] + offG[i]);
  if (alive[i]) {
    const int j = out.ptype[i] == HEPEM_GAMMA ? jG++ : jE++;
    CopyTrack(out, i, d, j);
    if (parent) parent[j] = i;
    if (origin) origin[j] = 0;
  }
  const uint32_t mask = smask[i];
  if (mask == 0u) return;
  const uint64_t sid = ((uint64_t)StreamHi(out, i) << 32) | (uint32_t)out.id[i];
  const uint32_t ctr = out.rng_ctr[i];
  for (int s = 0; s < 2; ++s) {
    if (!((mask >> s) & 1u)) continue;
    const int8_t pt = res.sec_ptype[s * n + i];
    const int j = pt == HEPEM_GAMMA ? jG++ : jE++;
    const uint64_t cid = Mix64(sid ^ Mix64(((uint64_t)ctr << 2) | (uint64_t)(s + 1)));
    d.ptype[j] = pt;
    d.id[j] = (int32_t)(uint32_t)cid;
    d.rng_stream[j] = (uint32_t)(cid >> 32);
    d.imc[j] = out.imc[i];
    d.ekin[j] = res.sec_ekin[s * n + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) d.dir[c * m + j] = res.sec_dir[(c * 2 + s) * n + i];
    d.safety[j] = out.safety[i];
    d.on_boundary[j] = out.on_boundary[i];
    // start-of-track state (G4HepEmTrack::ReSet, G4HepEmMSCTrackData::ReSet,
    // G4HepEmRandomEngine::DiscardGauss)
#pragma unroll
    for (int k = 0; k < 4; ++k) d.num_ia_left[k * m + j] = -1.0;
    d.msc_init_range[j] = 1.0e+21;
    d.msc_dyn_range_factor[j] = 0.04;
    d.msc_tlimit_min[j] = 1.0E-7;
    d.msc_first_step[j] = 1;
    d.rng_ctr[j] = 0u;
    d.rng_gauss[j] = 0.0;
    d.rng_has_gauss[j] = 0;
    if (parent) parent[j] = i;
    if (origin) origin[j] = (int8_t)(s + 1);
  }
}

bool TracksOk(const hepem_tracks* t) {
  return t && t->n >= 0 &&
         (t->n == 0 || (t->ptype && t->id && t->imc && t->ekin && t->dir && t->safety && t->on_boundary && t->num_ia_left &&
                        t->msc_init_range && t->msc_dyn_range_factor && t->msc_tlimit_min && t->msc_first_step && t->rng_ctr &&
                        t->rng_gauss && t->rng_has_gauss));
}

bool ResOk(const hepem_step_out* r) { return r && r->nsec && r->sec_ptype && r->sec_ekin && r->sec_dir; }

}  // namespace

struct hepem_stack {
  const hepem_ctx* ctx = nullptr;
  int cap = 0;
  int n = 0;                      // tracks of the current batch
  int cur = 0;                    // buffer holding it
  hepem_tracks buf[2] = {};
  CoupleCuts* cuts = nullptr;     // per couple
  int nmc = 0;
  uint8_t* alive = nullptr;       // per slot of a step's output
  uint8_t* smask = nullptr;
  uint32_t* offE = nullptr;
  uint32_t* offG = nullptr;
  uint32_t* bsumE = nullptr;      // per block
  uint32_t* bsumG = nullptr;
  uint32_t* total = nullptr;      // 2 words
};

namespace {
void FreeStack(hepem_stack* s) {
  FreeTracks(s->buf[0]);
  FreeTracks(s->buf[1]);
  void* ps[] = {s->cuts, s->alive, s->smask, s->offE, s->offG, s->bsumE, s->bsumG, s->total};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete s;
}
}  // namespace

extern "C" int hepem_stack_create(const hepem_ctx* ctx, int32_t capacity, hepem_stack** out) {
  if (!out) return HEPEM_ERR_INVALID;
  *out = nullptr;
  if (!ctx || capacity < 1) return HEPEM_ERR_INVALID;
  hepem_stack* s = new (std::nothrow) hepem_stack();
  if (!s) return HEPEM_ERR_ALLOC;
  s->ctx = ctx;
  s->cap = capacity;
  s->nmc = HepemTables64(ctx).nmc;
  std::vector<CoupleCuts> hc((size_t)(s->nmc > 0 ? s->nmc : 1));
  for (int k = 0; k < s->nmc; ++k) {
    double c[3];
    int apply = 0;
    if (HepemCoupleCuts(ctx, k, c, &apply) != HEPEM_OK) {
      FreeStack(s);
      return HEPEM_ERR_INVALID;
    }
    hc[(size_t)k] = CoupleCuts{c[0], c[1], c[2], apply};
  }
  const size_t cap = (size_t)capacity, nb = (cap + kB - 1) / kB;
  if (!AllocTracks(s->buf[0], capacity) || !AllocTracks(s->buf[1], capacity) ||
      hipMalloc(&s->cuts, hc.size() * sizeof(CoupleCuts)) != hipSuccess || hipMalloc(&s->alive, cap) != hipSuccess ||
      hipMalloc(&s->smask, cap) != hipSuccess || hipMalloc(&s->offE, cap * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&s->offG, cap * sizeof(uint32_t)) != hipSuccess || hipMalloc(&s->bsumE, nb * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&s->bsumG, nb * sizeof(uint32_t)) != hipSuccess || hipMalloc(&s->total, 2 * sizeof(uint32_t)) != hipSuccess) {
    FreeStack(s);
    return HEPEM_ERR_ALLOC;
  }
  if (hipMemcpy(s->cuts, hc.data(), hc.size() * sizeof(CoupleCuts), hipMemcpyHostToDevice) != hipSuccess) {
    FreeStack(s);
    return HEPEM_ERR_HIP;
  }
  *out = s;
  return HEPEM_OK;
}

extern "C" int hepem_stack_destroy(hepem_stack* s) {
  if (!s) return HEPEM_ERR_INVALID;
  FreeStack(s);
  return HEPEM_OK;
}

extern "C" int hepem_stack_tracks(const hepem_stack* s, hepem_tracks* view) {
  if (!s || !view) return HEPEM_ERR_INVALID;
  *view = s->buf[s->cur];
  view->n = s->n;
  return HEPEM_OK;
}

extern "C" int hepem_stack_push(hepem_stack* s, const hepem_tracks* t, void* stream) {
  if (!s || !TracksOk(t)) return HEPEM_ERR_INVALID;
  if (t->n == 0) return HEPEM_OK;
  if ((int64_t)s->n + t->n > s->cap) return HEPEM_ERR_CAPACITY;
  hepem_tracks a = s->buf[s->cur], d = s->buf[1 - s->cur];
  a.n = s->n;
  d.n = s->n + t->n;
  // the planar stride changes with the count: the whole batch is laid out
  // again in the other buffer
  hipLaunchKernelGGL(k_st_append, dim3((d.n + kB - 1) / kB), dim3(kB), 0, static_cast<hipStream_t>(stream), a, *t, d);
  if (hipGetLastError() != hipSuccess) return HEPEM_ERR_HIP;
  s->cur = 1 - s->cur;
  s->n = d.n;
  return HEPEM_OK;
}

extern "C" int hepem_stack_copy_out(const hepem_stack* s, hepem_tracks* dst, void* stream) {
  if (!s || !TracksOk(dst) || dst->n != s->n) return HEPEM_ERR_INVALID;
  if (s->n == 0) return HEPEM_OK;
  hepem_tracks a = s->buf[s->cur], none{};
  a.n = s->n;
  hipLaunchKernelGGL(k_st_append, dim3((s->n + kB - 1) / kB), dim3(kB), 0, static_cast<hipStream_t>(stream), a, none, *dst);
  return hipGetLastError() == hipSuccess ? HEPEM_OK : HEPEM_ERR_HIP;
}

extern "C" int hepem_stack_advance(hepem_stack* s, const hepem_tracks* out, const hepem_step_out* res, const uint8_t* keep,
                                   double* cut_edep, int32_t* parent, int8_t* origin, int32_t* n_next, void* stream) {
  if (!s || !TracksOk(out) || !n_next || (out->n > 0 && !ResOk(res))) return HEPEM_ERR_INVALID;
  if (out->n > s->cap) return HEPEM_ERR_CAPACITY;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = out->n;
  const hepem_tracks& target = s->buf[1 - s->cur];
  if (out->ptype == target.ptype) return HEPEM_ERR_INVALID;  // the step's output may not be the stack's other buffer
  if (n == 0) {
    s->n = 0;
    *n_next = 0;
    return HEPEM_OK;
  }
  const int nb = (n + kB - 1) / kB;
  hipLaunchKernelGGL(k_st_count, dim3(nb), dim3(kB), 0, st, *out, *res, keep, s->cuts, s->nmc, s->alive, s->smask, s->offE,
                     s->offG, s->bsumE, s->bsumG, cut_edep);
  hipLaunchKernelGGL(k_st_scan, dim3(1), dim3(kScanB), 0, st, s->bsumE, s->bsumG, nb, s->total);
  hepem_tracks d = target;
  d.n = s->cap;  // set by the kernel to the new count
  hipLaunchKernelGGL(k_st_scatter, dim3(nb), dim3(kB), 0, st, *out, *res, s->alive, s->smask, s->offE, s->offG, s->bsumE,
                     s->bsumG, s->total, d, s->cap, parent, origin);
  if (hipGetLastError() != hipSuccess) return HEPEM_ERR_HIP;
  uint32_t total[2] = {0u, 0u};
  if (hipMemcpyAsync(total, s->total, sizeof(total), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return HEPEM_ERR_HIP;
  *n_next = (int32_t)total[0];
  if (total[0] > (uint32_t)s->cap) return HEPEM_ERR_CAPACITY;  // nothing was written: the batch stays
  s->cur = 1 - s->cur;
  s->n = (int)total[0];
  return HEPEM_OK;
}
