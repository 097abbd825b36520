// One-track drop-in (include/g4hepem_amd_onetrack.hpp) on the GPU: a
// TrackingManager-style loop over single tracks through
// G4HepEmElectronManager / G4HepEmGammaManager HowFar, a geometry cut and
// Perform, against the same tracks stepped as one mixed batch through the
// C-ABI (hepem_howfar, the same cut, hepem_perform), bit for bit: the HowFar
// results, the post-step state, the deposit and the secondaries. Run by
// tests/test_gpu_cpp.py. Prints one line per check; exit status 0 when all pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "g4hepem_amd_onetrack.hpp"
#include "g4hepem_amd_tables.h"

using namespace g4hepem_amd;
namespace ot = g4hepem_amd::onetrack;

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
bool Same(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

struct Tracks {
  std::vector<int8_t> pt;
  std::vector<int32_t> id, imc;
  std::vector<double> e, dir, saf;
};

// e-:e+:gamma = 4:1:5 by index, energy log-uniform in [10 keV, 1 TeV] from
// an LCG, isotropic, not on a boundary
Tracks Make(int n, int ncouples) {
  Tracks t;
  t.pt.resize(n);
  t.id.resize(n);
  t.imc.resize(n);
  t.e.resize(n);
  t.dir.resize(3 * (size_t)n);
  t.saf.resize(n);
  uint64_t s = 0x2545F4914F6CDD1Dull;
  auto u = [&s]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return ((double)(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < n; ++i) {
    t.pt[i] = (int8_t)(i % 10 < 4 ? 0 : i % 10 < 5 ? 1 : 2);
    t.id[i] = 1000 + i;
    t.imc[i] = (int32_t)(u() * ncouples);
    t.e[i] = std::exp(std::log(1e-2) + u() * (std::log(1e6) - std::log(1e-2)));
    const double c = 2.0 * u() - 1.0, p = 6.283185307179586 * u(), sn = std::sqrt((1.0 - c) * (1.0 + c));
    t.dir[i] = sn * std::cos(p);
    t.dir[n + i] = sn * std::sin(p);
    t.dir[2 * (size_t)n + i] = c;
    t.saf[i] = std::exp(std::log(1e-3) + u() * std::log(1e4));
  }
  return t;
}

// the geometry of both paths: every third track's step ends on a boundary at
// half its proposed length
bool Cut(int i) { return i % 3 == 0; }

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
    const uint64_t seed = 0x0E7AC;
    const int n = 3000;
    const Tracks h = Make(n, 4);
    // ---- the batch through the C-ABI
    Context ctx(data, pars);
    TrackBatch b(n);
    HowFarState hs(n);
    StepResults r(n);
    WorkBuffer w(n);
    hepem_tracks& t = b.View();
    Upload(t.ptype, h.pt);
    Upload(t.id, h.id);
    Upload(t.imc, h.imc);
    Upload(t.ekin, h.e);
    Upload(t.dir, h.dir);
    Upload(t.safety, h.saf);
    Upload(t.on_boundary, std::vector<uint8_t>(n, 0));
    b.ResetPhysicsState();
    Stepper::HowFar(ctx, seed, b, hs, w);
    std::vector<double> gstep = Download(hs.View().geom_step, n);
    const std::vector<double> gstepHowFar = gstep, mfp = Download(hs.View().mfp, 4 * (size_t)n);
    const std::vector<int32_t> winHowFar = Download(hs.View().winner, n);
    std::vector<uint8_t> onb(n);
    for (int i = 0; i < n; ++i) {
      onb[i] = Cut(i) ? 1 : 0;
      if (Cut(i)) gstep[i] *= 0.5;
    }
    Upload(hs.View().geom_step, gstep);
    Upload(t.on_boundary, onb);
    Stepper::Perform(ctx, seed, b, hs, r, w);
    const std::vector<double> ekin = Download(t.ekin, n), dir = Download(t.dir, 3 * (size_t)n),
                              nial = Download(t.num_ia_left, 4 * (size_t)n), edep = Download(r.View().edep, n),
                              secE = Download(r.View().sec_ekin, 2 * (size_t)n),
                              secD = Download(r.View().sec_dir, 6 * (size_t)n);
    const std::vector<uint32_t> ctr = Download(t.rng_ctr, n);
    const std::vector<int32_t> nsec = Download(r.View().nsec, n);
    const std::vector<int8_t> secP = Download(r.View().sec_ptype, 2 * (size_t)n);

    // ---- the same tracks one at a time, as a tracking manager drives them
    ot::G4HepEmTLData tl;
    tl.SetSeed(seed);
    bool howfarOk = true, stateOk = true, depOk = true, secOk = true;
    int nsecTotal = 0;
    for (int i = 0; i < n; ++i) {
      const bool isGamma = h.pt[i] == 2;
      ot::G4HepEmTrack* tr;
      if (isGamma) {
        tl.GetPrimaryGammaTrack()->ReSet();
        tr = tl.GetPrimaryGammaTrack()->GetTrack();
      } else {
        tl.GetPrimaryElectronTrack()->ReSet();
        tr = tl.GetPrimaryElectronTrack()->GetTrack();
        tr->SetCharge(h.pt[i] == 0 ? -1.0 : 1.0);
      }
      tr->SetID(h.id[i]);
      tr->SetMCIndex(h.imc[i]);
      tr->SetEKin(h.e[i]);
      tr->SetDirection(h.dir[i], h.dir[n + i], h.dir[2 * (size_t)n + i]);
      tr->SetSafety(h.saf[i]);
      tr->SetOnBoundary(false);
      if (isGamma) ot::G4HepEmGammaManager::HowFar(data, pars, &tl);
      else ot::G4HepEmElectronManager::HowFar(data, pars, &tl);
      howfarOk = howfarOk && Same(tr->GetGStepLength(), gstepHowFar[i]) && tr->GetWinnerProcessIndex() == winHowFar[i];
      for (int k = 0; k < (isGamma ? 1 : 4); ++k) howfarOk = howfarOk && Same(tr->GetMFP(k), mfp[(size_t)k * n + i]);
      // the geometry
      if (Cut(i)) tr->SetGStepLength(tr->GetGStepLength() * 0.5);
      tr->SetOnBoundary(Cut(i));
      tl.ResetNumSecondaryElectronTrack();
      tl.ResetNumSecondaryGammaTrack();
      if (isGamma) ot::G4HepEmGammaManager::Perform(data, pars, &tl);
      else ot::G4HepEmElectronManager::Perform(data, pars, &tl);
      stateOk = stateOk && Same(tr->GetEKin(), ekin[i]) && tr->GetRngCounter() == ctr[i];
      for (int c = 0; c < 3; ++c) stateOk = stateOk && Same(tr->GetDirection()[c], dir[(size_t)c * n + i]);
      for (int k = 0; k < 4; ++k) stateOk = stateOk && Same(tr->GetNumIALeft(k), nial[(size_t)k * n + i]);
      depOk = depOk && Same(tr->GetEnergyDeposit(), edep[i]);
      // secondaries: the C-ABI's slots hold electrons first, then gammas,
      // as the TLData's two lists read one after the other
      const int ne = (int)tl.GetNumSecondaryElectronTrack(), ng = (int)tl.GetNumSecondaryGammaTrack();
      secOk = secOk && ne + ng == nsec[i];
      for (int s = 0; s < nsec[i] && s < 2 && secOk; ++s) {
        ot::G4HepEmTrack* st = s < ne ? tl.GetSecondaryElectronTrack(s)->GetTrack() : tl.GetSecondaryGammaTrack(s - ne)->GetTrack();
        const int8_t p = secP[(size_t)s * n + i];
        const double charge = p == 2 ? 0.0 : p == 0 ? -1.0 : 1.0;
        secOk = st->GetCharge() == charge && Same(st->GetEKin(), secE[(size_t)s * n + i]) && st->GetParentID() == h.id[i];
        for (int c = 0; c < 3; ++c) secOk = secOk && Same(st->GetDirection()[c], secD[(size_t)(c * 2 + s) * n + i]);
      }
      nsecTotal += nsec[i];
    }
    std::printf("%d tracks one at a time, %d secondaries\n", n, nsecTotal);
    Expect(howfarOk, "HowFar: geometrical step, MFPs and winner equal the batch's");
    Expect(stateOk, "Perform: post-step energy, direction, interaction lengths and draw count equal the batch's");
    Expect(depOk, "Perform: energy deposit equals the batch's");
    Expect(secOk && nsecTotal > 100, "Perform: secondaries (kind, energy, direction, parent) equal the batch's slots");
    // the device image is made once per (data, pars) pair
    Expect(ot::detail::ContextFor(data, pars) == ot::detail::ContextFor(data, pars), "one device image per table pair");
    // an error surfaces as an exception with the C-ABI status
    bool threw = false;
    try {
      ot::G4HepEmTLData bad;
      bad.GetPrimaryGammaTrack()->ReSet();
      bad.GetPrimaryGammaTrack()->GetTrack()->SetEKin(1.0);
      bad.GetPrimaryGammaTrack()->GetTrack()->SetMCIndex(0);
      ot::G4HepEmGammaManager::HowFar(nullptr, pars, &bad);
    } catch (const Error& e) {
      threw = e.Status() == HEPEM_ERR_INVALID;
    }
    Expect(threw, "a null table pointer raises Error(HEPEM_ERR_INVALID)");
  } catch (const Error& e) {
    std::printf("FAIL: exception %s\n", e.what());
    rc = 1;
  }
  hepem_tables_free(data, pars);
  if (gFailures == 0 && rc == 0) std::printf("ALL OK\n");
  return gFailures == 0 ? rc : 1;
}
