// C++ host layer (include/g4hepem_amd.hpp) on the GPU: the per-kind managers,
// the mixed-batch HowFar/Perform and the fused step must agree bit for bit,
// and errors surface as g4hepem_amd::Error. Run by tests/test_gpu_cpp.py.
// Prints one line per check; exit status 0 when all pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "g4hepem_amd.hpp"
#include "g4hepem_amd_tables.h"

using namespace g4hepem_amd;

namespace {

int gFailures = 0;
void Expect(bool ok, const char* what) {
  std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
  std::fflush(stdout);
  if (!ok) ++gFailures;
}

template <typename T> void Upload(T* d, const std::vector<T>& h) {
  CheckHip("hipMemcpy", hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}
template <typename T> std::vector<T> Download(const T* d, size_t n) {
  std::vector<T> h(n);
  CheckHip("hipMemcpy", hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

// fresh tracks: particle kind by `kind` (-1: e-:e+:gamma = 4:1:5 by index),
// energy log-uniform in [10 keV, 1 TeV] from an LCG, isotropic directions
void Fill(TrackBatch& b, int kind, int ncouples, uint32_t first_id) {
  const int n = b.Size();
  std::vector<int8_t> pt(n);
  std::vector<int32_t> id(n), imc(n);
  std::vector<double> e(n), dir(3 * (size_t)n), saf(n);
  std::vector<uint8_t> onb(n);
  uint64_t s = 0x9E3779B97F4A7C15ull ^ first_id;
  auto u = [&s]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return ((double)(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < n; ++i) {
    pt[i] = (int8_t)(kind >= 0 ? kind : (i % 10 < 4 ? 0 : i % 10 < 5 ? 1 : 2));
    id[i] = (int32_t)(first_id + (uint32_t)i);
    imc[i] = (int32_t)(u() * ncouples);
    e[i] = std::exp(std::log(1e-2) + u() * (std::log(1e6) - std::log(1e-2)));
    const double c = 2.0 * u() - 1.0, p = 6.283185307179586 * u(), sn = std::sqrt((1.0 - c) * (1.0 + c));
    dir[i] = sn * std::cos(p);
    dir[n + i] = sn * std::sin(p);
    dir[2 * (size_t)n + i] = c;
    saf[i] = std::exp(std::log(1e-3) + u() * std::log(1e4));
    onb[i] = (uint8_t)(i % 7 == 3);
  }
  hepem_tracks& t = b.View();
  Upload(t.ptype, pt);
  Upload(t.id, id);
  Upload(t.imc, imc);
  Upload(t.ekin, e);
  Upload(t.dir, dir);
  Upload(t.safety, saf);
  Upload(t.on_boundary, onb);
  b.ResetPhysicsState();
}

void Copy(TrackBatch& dst, TrackBatch& src) {
  const hepem_tracks &a = src.View(), &b = dst.View();
  const size_t n = (size_t)a.n;
  auto cp = [](void* d, const void* s, size_t bytes) {
    CheckHip("hipMemcpy", hipMemcpy(d, s, bytes, hipMemcpyDeviceToDevice));
  };
  cp(b.ptype, a.ptype, n);
  cp(b.id, a.id, 4 * n);
  cp(b.imc, a.imc, 4 * n);
  cp(b.ekin, a.ekin, 8 * n);
  cp(b.dir, a.dir, 24 * n);
  cp(b.safety, a.safety, 8 * n);
  cp(b.on_boundary, a.on_boundary, n);
  cp(b.num_ia_left, a.num_ia_left, 32 * n);
  cp(b.msc_init_range, a.msc_init_range, 8 * n);
  cp(b.msc_dyn_range_factor, a.msc_dyn_range_factor, 8 * n);
  cp(b.msc_tlimit_min, a.msc_tlimit_min, 8 * n);
  cp(b.msc_first_step, a.msc_first_step, n);
  cp(b.rng_ctr, a.rng_ctr, 4 * n);
  cp(b.rng_gauss, a.rng_gauss, 8 * n);
  cp(b.rng_has_gauss, a.rng_has_gauss, n);
}

template <typename T> bool SameBytes(const T* a, const T* b, size_t n) {
  const std::vector<T> x = Download(a, n), y = Download(b, n);
  return std::memcmp(x.data(), y.data(), n * sizeof(T)) == 0;
}

// secondary slots at or beyond nsec are not written, so only valid ones compare
bool SameSecondaries(hepem_step_out& x, hepem_step_out& y, size_t n) {
  const std::vector<int32_t> ns = Download(x.nsec, n);
  const std::vector<int8_t> pa = Download(x.sec_ptype, 2 * n), pb = Download(y.sec_ptype, 2 * n);
  const std::vector<double> ea = Download(x.sec_ekin, 2 * n), eb = Download(y.sec_ekin, 2 * n);
  const std::vector<double> da = Download(x.sec_dir, 6 * n), db = Download(y.sec_dir, 6 * n);
  for (size_t i = 0; i < n; ++i)
    for (int s = 0; s < ns[i] && s < 2; ++s) {
      if (pa[s * n + i] != pb[s * n + i] || std::memcmp(&ea[s * n + i], &eb[s * n + i], sizeof(double)) != 0) return false;
      for (int c = 0; c < 3; ++c)
        if (std::memcmp(&da[(c * 2 + s) * n + i], &db[(c * 2 + s) * n + i], sizeof(double)) != 0) return false;
    }
  return true;
}

bool SameResults(StepResults& a, StepResults& b, size_t n) {
  hepem_step_out &x = a.View(), &y = b.View();
  return SameBytes(x.true_step, y.true_step, n) && SameBytes(x.geom_step, y.geom_step, n) && SameBytes(x.edep, y.edep, n) &&
         SameBytes(x.displacement, y.displacement, 3 * n) && SameBytes(x.winner, y.winner, n) &&
         SameBytes(x.stopped, y.stopped, n) && SameBytes(x.nsec, y.nsec, n) && SameSecondaries(x, y, n);
}
bool SameState(TrackBatch& a, TrackBatch& b, size_t n) {
  hepem_tracks &x = a.View(), &y = b.View();
  return SameBytes(x.ekin, y.ekin, n) && SameBytes(x.dir, y.dir, 3 * n) && SameBytes(x.num_ia_left, y.num_ia_left, 4 * n) &&
         SameBytes(x.rng_ctr, y.rng_ctr, n) && SameBytes(x.rng_gauss, y.rng_gauss, n);
}

}  // namespace

int main() {
  const char* names[] = {"G4_PbWO4", "G4_Pb", "G4_lAr", "G4_Si"};
  G4HepEmData* data = nullptr;
  G4HepEmParameters* pars = nullptr;
  if (hepem_tables_build(names, 4, 0.7, &data, &pars) != 0) {
    std::printf("FAIL: table build\n");
    return 1;
  }
  int rc = 0;
  try {
    Context ctx(data, pars);
    const uint64_t seed = 0xC0DE;
    for (int kind : {-1, 0, 2}) {
      const int n = 40000;
      TrackBatch in(n), fused(n), split(n);
      Fill(in, kind, 4, 1000u + 100000u * (uint32_t)(kind + 1));
      Copy(split, in);
      StepResults rf(n), rs(n);
      WorkBuffer w(n);
      HowFarState hs(n);
      Stepper::Step(ctx, seed, in, fused, rf, w);
      if (kind == 0) {
        ElectronManager::HowFar(ctx, seed, split, hs, w);
        ElectronManager::Perform(ctx, seed, split, hs, rs, w);
      } else if (kind == 2) {
        GammaManager::HowFar(ctx, seed, split, hs, w);
        GammaManager::Perform(ctx, seed, split, hs, rs, w);
      } else {
        Stepper::HowFar(ctx, seed, split, hs, w);
        Stepper::Perform(ctx, seed, split, hs, rs, w);
      }
      CheckHip("hipDeviceSynchronize", hipDeviceSynchronize());
      const char* label = kind == 0 ? "ElectronManager HowFar+Perform == fused step (e- batch)"
                          : kind == 2 ? "GammaManager HowFar+Perform == fused step (gamma batch)"
                                      : "Stepper HowFar+Perform == fused step (mixed batch)";
      Expect(SameResults(rf, rs, n) && SameState(fused, split, n), label);
      Expect(w.Rejected() == 0, "no rejected tracks");
    }
    // a gamma in an electron batch is rejected, not stepped
    {
      TrackBatch t(100);
      Fill(t, -1, 4, 7u);
      HowFarState hs(100);
      StepResults r(100);
      WorkBuffer w(100);
      ElectronManager::HowFar(ctx, seed, t, hs, w);
      ElectronManager::Perform(ctx, seed, t, hs, r, w);
      const std::vector<int32_t> win = Download(r.View().winner, 100);
      int bad = 0;
      for (int i = 0; i < 100; ++i) bad += ((i % 10 >= 5) != (win[i] == -99));
      Expect(bad == 0 && w.Rejected() == 50, "gammas in an electron batch are rejected");
    }
    // errors are exceptions with the C-ABI status
    bool threw = false;
    try {
      Context bad(data, pars, static_cast<Precision>(7));
    } catch (const Error& e) {
      threw = e.Status() == HEPEM_ERR_INVALID;
    }
    Expect(threw, "invalid precision throws Error(HEPEM_ERR_INVALID)");
  } catch (const Error& e) {
    std::printf("FAIL: exception %s\n", e.what());
    rc = 1;
  }
  hepem_tables_free(data, pars);
  if (gFailures) rc = 1;
  std::printf("%s\n", rc == 0 ? "ALL OK" : "FAILURES");
  return rc;
}
