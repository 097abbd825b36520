// hepem_capi.hip -- the C-ABI of the library (declared in include/g4hepem_amd.h):
// context creation (one device image of every table, FP64 and FP32 copies),
// and the launch wrappers of the lookup and stepping kernels.
#include "hepem_kernels.hh"
#include "hepem_internal.hh"
#include "tables/hepem_validate.hh"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

using namespace hepem;

#define HIP_OK(call)                                    \
  do {                                                  \
    if ((call) != hipSuccess) return HEPEM_ERR_HIP;     \
  } while (0)

namespace {

// Host staging of the device image: arrays appended at 256-byte alignment;
// pointer fields of the DevTables structs are patched after the upload.
struct Packer {
  std::vector<uint8_t> buf;
  std::vector<std::pair<const void**, size_t>> fix;
  template <typename T> void add(const T*& field, const T* src, size_t n) {
    size_t off = (buf.size() + 255) & ~size_t(255);
    buf.resize(off + (n ? n : 1) * sizeof(T), 0);
    if (n && src) std::memcpy(buf.data() + off, src, n * sizeof(T));
    fix.emplace_back(reinterpret_cast<const void**>(&field), off);
  }
  template <typename T> void add(const T*& field, const std::vector<T>& v) { add(field, v.data(), v.size()); }
  // table values converted to the form's precision R
  template <typename R> void addR(const R*& field, const double* src, size_t n) {
    std::vector<R> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (R)src[i];
    add(field, v);
  }
  // a table whose abscissas are also read in double: in the FP64 form both
  // fields alias one array, in the FP32 form a double copy is added
  void addX(const double*& field, const double*& field64, const double* src, size_t n) {
    add(field, src, n);
    fix.emplace_back(reinterpret_cast<const void**>(&field64), fix.back().second);
  }
  void addX(const float*& field, const double*& field64, const double* src, size_t n) {
    addR(field, src, n);
    add(field64, src, n);
  }
};

// Search guide of the SB sampling tables: for every (element, energy) row,
// entry g = largest kappa index k whose cumulative value (in precision R, as
// the kernels compare it) is <= g / kSBGuide. Rows are stored per element in
// table order; sbGuideStartZ[Z] is the first row of element Z.
template <typename R> void PackSBGuide(Packer& P, DevTables<R>& T, const G4HepEmSBTableData* sb) {
  std::vector<uint8_t> guide;
  std::vector<int> startZ(121, -1);
  const int nk = sb->fNumKappa;
  int rows = 0;
  for (int z = 0; z < 121; ++z) {
    const int s = sb->fSBTablesStartPerZ[z];
    if (s < 0) continue;
    const double* d = sb->fSBTableData + s;
    const int minE = (int)d[1], maxE = (int)d[2], ncut = (int)d[3];
    startZ[z] = rows;
    for (int ie = 0; ie <= maxE - minE; ++ie) {
      const double* cum = d + 4 + (size_t)ie * (ncut + 3 * nk) + ncut;
      for (int g = 0; g < kSBGuide; ++g) {
        const R t = (R)((double)g / kSBGuide);
        int k = 0;
        while (k + 1 < nk && (R)cum[3 * (k + 1)] <= t) ++k;
        guide.push_back((uint8_t)k);
      }
      ++rows;
    }
  }
  P.add(T.sbGuide, guide);
  P.add(T.sbGuideStartZ, startZ);
}

template <typename R> void PackEl(Packer& P, DevEl<R>& E, const G4HepEmElectronData* e) {
  const int N = e->fELossEnergyGridSize;
  E.N = N;
  E.lossLogMin = (R)e->fELossLogMinEkin;
  E.lossILD = (R)e->fELossEILDelta;
  P.addX(E.lossGrid, E.lossGrid64, e->fELossEnergyGrid, N);
  E.lossEmin = N > 0 ? e->fELossEnergyGrid[0] : 0.0;
  E.lossEmax = N > 0 ? e->fELossEnergyGrid[N - 1] : 0.0;
  P.addX(E.lossData, E.lossData64, e->fELossData, (size_t)5 * N * e->fNumMatCuts);
  // the inverse-range search may start from a bin hint only where the search
  // result cannot depend on the path: non-decreasing range columns
  E.rangeMonotone = 1;
  for (int mc = 0; mc < e->fNumMatCuts && E.rangeMonotone; ++mc) {
    const double* r = e->fELossData + (size_t)5 * N * mc;
    for (int k = 0; k + 1 < N; ++k)
      if (!(r[2 * k] <= r[2 * (k + 1)])) {
        E.rangeMonotone = 0;
        break;
      }
  }
  P.add(E.xsStart, e->fResMacXSecStartIndexPerMatCut, e->fNumMatCuts);
  P.addX(E.xsData, E.xsData64, e->fResMacXSecData, e->fResMacXSecNumData);
  E.xsN = e->fResMacXSecNumData;
  {
    // block header: [0] triples, [1] energy of the largest value, [2] largest
    // value, [3] log of the first energy, [4] 1/log-spacing, then the
    // [energy, value, second derivative] triples from [5]
    std::vector<XsBlock<R>> blocks((size_t)2 * e->fNumMatCuts);
    const double* d = e->fResMacXSecData;
    for (int mc = 0; mc < e->fNumMatCuts; ++mc) {
      const int iIoni = e->fResMacXSecStartIndexPerMatCut[mc];
      const int nIoni = (int)d[iIoni];
      for (int k = 0; k < 2; ++k) {
        const int s0 = k == 0 ? iIoni : iIoni + 3 * nIoni + 5;
        XsBlock<R>& b = blocks[2 * (size_t)mc + k];
        b.n = (int)d[s0];
        b.minE = d[s0 + 5];
        b.lastE = b.n > 0 ? d[s0 + 5 + 3 * (b.n - 1)] : b.minE;
        b.maxE = d[s0 + 1];
        b.start = s0 + 5;
        b.maxV = (R)d[s0 + 2];
        b.logMin = (R)d[s0 + 3];
        b.ild = (R)d[s0 + 4];
      }
    }
    P.add(E.xsBlock, blocks);
  }
  E.nucN = e->fENucEnergyGridSize;
  E.nucLogMin = (R)e->fENucLogMinEkin;
  E.nucILD = (R)e->fENucEILDelta;
  P.addX(E.nucGrid, E.nucGrid64, e->fENucEnergyGrid, e->fENucEnergyGridSize);
  E.nucEmin = e->fENucEnergyGridSize > 0 ? e->fENucEnergyGrid[0] : 0.0;
  E.nucEmax = e->fENucEnergyGridSize > 0 ? e->fENucEnergyGrid[e->fENucEnergyGridSize - 1] : 0.0;
  P.addR(E.nucData, e->fENucMacXsecData, (size_t)2 * e->fENucEnergyGridSize * e->fNumMaterials);
  P.addR(E.tr1, e->fTr1MacXSecData, (size_t)2 * N * e->fNumMaterials);
  P.add(E.selSBStart, e->fElemSelectorBremSBStartIndexPerMatCut, e->fNumMatCuts);
  P.addX(E.selSB, E.selSB64, e->fElemSelectorBremSBData, e->fElemSelectorBremSBNumData);
  P.add(E.selRBStart, e->fElemSelectorBremRBStartIndexPerMatCut, e->fNumMatCuts);
  P.addX(E.selRB, E.selRB64, e->fElemSelectorBremRBData, e->fElemSelectorBremRBNumData);
}

template <typename R> void PackTables(Packer& P, DevTables<R>& T, const G4HepEmData* d, const G4HepEmParameters* p) {
  PackEl(P, T.el[0], d->fTheElectronData);
  PackEl(P, T.el[1], d->fThePositronData);
  const G4HepEmMatCutData* mc = d->fTheMatCutData;
  const int nmc = mc->fNumMatCutData;
  T.nmc = nmc;
  std::vector<double> eCut(nmc), gCut(nmc), lgCut(nmc);
  std::vector<int> mMat(nmc), mReg(nmc);
  for (int i = 0; i < nmc; ++i) {
    eCut[i] = mc->fMatCutData[i].fSecElProdCutE;
    gCut[i] = mc->fMatCutData[i].fSecGamProdCutE;
    lgCut[i] = mc->fMatCutData[i].fLogSecGamCutE;
    mMat[i] = mc->fMatCutData[i].fHepEmMatIndex;
    mReg[i] = mc->fMatCutData[i].fG4RegionIndex;
  }
  P.add(T.mcElCut, eCut);
  P.add(T.mcGamCut, gCut);
  P.add(T.mcLogGamCut, lgCut);
  P.add(T.mcMat, mMat);
  P.add(T.mcReg, mReg);
  const G4HepEmMaterialData* md = d->fTheMaterialData;
  const int nmat = md->fNumMaterialData;
  T.nmat = nmat;
  std::vector<int> nEl(nmat), elStart(nmat), elZ, sStart(nmat), sN(nmat);
  std::vector<double> nat, densCor(nmat), elDens(nmat), radLen(nmat), meanExc(nmat), zeff(nmat), zeff23(nmat), zeffSqrt(nmat),
      umscPar(nmat), stepMin(2 * nmat), tail(4 * nmat), theta(2 * nmat), sE, sC;
  for (int i = 0; i < nmat; ++i) {
    const G4HepEmMatData& m = md->fMaterialData[i];
    nEl[i] = m.fNumOfElement;
    elStart[i] = (int)elZ.size();
    for (int k = 0; k < m.fNumOfElement; ++k) {
      elZ.push_back(m.fElementVect[k]);
      nat.push_back(m.fNumOfAtomsPerVolumeVect[k]);
    }
    densCor[i] = m.fDensityCorFactor;
    elDens[i] = m.fElectronDensity;
    radLen[i] = m.fRadiationLength;
    meanExc[i] = m.fMeanExEnergy;
    zeff[i] = m.fZeff;
    zeff23[i] = m.fZeff23;
    zeffSqrt[i] = m.fZeffSqrt;
    umscPar[i] = m.fUMSCPar;
    for (int k = 0; k < 2; ++k) stepMin[2 * i + k] = m.fUMSCStepMinPars[k];
    for (int k = 0; k < 4; ++k) tail[4 * i + k] = m.fUMSCTailCoeff[k];
    for (int k = 0; k < 2; ++k) theta[2 * i + k] = m.fUMSCThetaCoeff[k];
    sStart[i] = (int)sE.size();
    sN[i] = m.fNumOfSandiaIntervals;
    for (int k = 0; k < m.fNumOfSandiaIntervals; ++k) sE.push_back(m.fSandiaEnergies[k]);
    for (int k = 0; k < 4 * m.fNumOfSandiaIntervals; ++k) sC.push_back(m.fSandiaCoefficients[k]);
  }
  P.add(T.matNElem, nEl);
  P.add(T.matElemStart, elStart);
  P.add(T.matElemZ, elZ);
  P.add(T.matElemNat, nat);
  P.add(T.matDensCor, densCor);
  P.add(T.matElDens, elDens);
  P.add(T.matRadLen, radLen);
  P.add(T.matMeanExc, meanExc);
  P.add(T.matZeff, zeff);
  P.add(T.matZeff23, zeff23);
  P.add(T.matZeffSqrt, zeffSqrt);
  P.add(T.matUMSCPar, umscPar);
  P.add(T.matStepMin, stepMin);
  P.add(T.matTail, tail);
  P.add(T.matTheta, theta);
  P.add(T.matSandiaStart, sStart);
  P.add(T.matSandiaN, sN);
  P.add(T.matSandiaE, sE);
  P.add(T.matSandiaC, sC);
  const G4HepEmElementData* ed = d->fTheElementData;
  const int nz = ed->fMaxZet + 1;
  T.maxZ = ed->fMaxZet;
  std::vector<double> z13(nz), z23(nz), coul(nz), logZ(nz), zf1(nz), dLow(nz), dHigh(nz), ils1(nz), ils1c(nz), kb(nz), eE, eC;
  std::vector<int> eS(nz), eN(nz);
  for (int z = 0; z < nz; ++z) {
    const G4HepEmElemData& e = ed->fElementData[z];
    z13[z] = e.fZet13;
    z23[z] = e.fZet23;
    coul[z] = e.fCoulomb;
    logZ[z] = e.fLogZ;
    zf1[z] = e.fZFactor1;
    dLow[z] = e.fDeltaMaxLow;
    dHigh[z] = e.fDeltaMaxHigh;
    ils1[z] = e.fILVarS1;
    ils1c[z] = e.fILVarS1Cond;
    kb[z] = e.fKShellBindingEnergy;
    eS[z] = (int)eE.size();
    const int ni = e.fZet >= 0.0 ? e.fNumOfSandiaIntervals : 0;
    eN[z] = ni;
    for (int k = 0; k < ni; ++k) eE.push_back(e.fSandiaEnergies[k]);
    for (int k = 0; k < 4 * ni; ++k) eC.push_back(e.fSandiaCoefficients[k]);
  }
  P.add(T.elZ13, z13);
  P.add(T.elZ23, z23);
  P.add(T.elCoulomb, coul);
  P.add(T.elLogZ, logZ);
  P.add(T.elZF1, zf1);
  P.add(T.elDMaxLow, dLow);
  P.add(T.elDMaxHigh, dHigh);
  P.add(T.elILS1, ils1);
  P.add(T.elILS1Cond, ils1c);
  P.add(T.elKBind, kb);
  P.add(T.elSandiaStart, eS);
  P.add(T.elSandiaN, eN);
  P.add(T.elSandiaE, eE);
  P.add(T.elSandiaC, eC);
  const G4HepEmSBTableData* sb = d->fTheSBTableData;
  P.addR(T.sbData, sb->fSBTableData, sb->fNumSBTableData);
  P.add(T.sbStartZ, sb->fSBTablesStartPerZ, 121);
  P.add(T.sbGCutStart, sb->fGammaCutIndxStartIndexPerMC, sb->fNumHepEmMatCuts);
  P.add(T.sbGCutIdx, sb->fGammaCutIndices, sb->fNumElemsInMatCuts);
  P.add(T.sbElE, sb->fElEnergyVect, 65);
  P.add(T.sbLElE, sb->fLElEnergyVect, 65);
  P.addR(T.sbLKappa, sb->fLKappaVect, 54);
  T.sbLogMinE = sb->fLogMinElEnergy;
  T.sbILDE = sb->fILDeltaElEnergy;
  T.sbNumKappa = sb->fNumKappa;
  PackSBGuide<R>(P, T, sb);
  const G4HepEmGammaData* g = d->fTheGammaData;
  P.addX(T.gmData, T.gmData64, g->fMacXsecData, (size_t)g->fNumMaterials * g->fDataPerMat);
  T.gmPerMat = g->fDataPerMat;
  T.gmOff1 = g->fNumData0;
  T.gmOff2 = g->fNumData0 + g->fNumData1;
  T.gmN0 = g->fEGridSize0;
  T.gmN1 = g->fEGridSize1;
  T.gmN2 = g->fEGridSize2;
  T.gEMax0 = g->fEMax0;
  T.gEMax1 = g->fEMax1;
  T.gLogEMin0 = (R)g->fLogEMin0;
  T.gILD0 = (R)g->fEILDelta0;
  T.gLogEMin1 = (R)g->fLogEMin1;
  T.gILD1 = (R)g->fEILDelta1;
  T.gLogEMin2 = (R)g->fLogEMin2;
  T.gILD2 = (R)g->fEILDelta2;
  P.add(T.convSelStart, g->fElemSelectorConvStartIndexPerMat, g->fNumMaterials);
  P.addR(T.convSelData, g->fElemSelectorConvData, g->fElemSelectorConvNumData);
  P.addX(T.convSelGrid, T.convSelGrid64, g->fElemSelectorConvEgrid, g->fElemSelectorConvEgridSize);
  T.convSelN = g->fElemSelectorConvEgridSize;
  T.convSelLogMin = (R)g->fElemSelectorConvLogMinEkin;
  T.convSelILD = (R)g->fElemSelectorConvEILDelta;
  T.elTrackCut = p->fElectronTrackingCut;
  T.gamTrackCut = p->fGammaTrackingCut;
  T.minLossE = p->fMinLossTableEnergy;
  T.bremModelLim = p->fElectronBremModelLim;
  T.posCor = p->fIsMSCPositronCor ? 1 : 0;
  T.displace = p->fIsMSCDisplacement ? 1 : 0;
  const int nreg = p->fNumRegions;
  std::vector<double> fr(nreg), dr(nreg), ll(nreg), rf(nreg), sf(nreg);
  std::vector<int> mn(nreg), fl(nreg), ms(nreg);
  for (int r = 0; r < nreg; ++r) {
    const G4HepEmRegionParmeters& rp = p->fParametersPerRegion[r];
    fr[r] = rp.fFinalRange;
    dr[r] = rp.fDRoverRange;
    ll[r] = rp.fLinELossLimit;
    rf[r] = rp.fMSCRangeFactor;
    sf[r] = rp.fMSCSafetyFactor;
    mn[r] = rp.fIsMSCMinimalStepLimit ? 1 : 0;
    fl[r] = rp.fIsELossFluctuation ? 1 : 0;
    ms[r] = rp.fIsMultipleStepsInMSCTrans ? 1 : 0;
  }
  P.add(T.regFinalRange, fr);
  P.add(T.regDRover, dr);
  P.add(T.regLinLim, ll);
  P.add(T.regRangeFac, rf);
  P.add(T.regSafetyFac, sf);
  P.add(T.regMinimal, mn);
  P.add(T.regFluct, fl);
  P.add(T.regMultiStep, ms);
}

bool DataComplete(const G4HepEmData* d, const G4HepEmParameters* p) {
  if (!d || !p || !d->fTheMatCutData || !d->fTheMaterialData || !d->fTheElementData || !d->fTheElectronData ||
      !d->fThePositronData || !d->fTheSBTableData || !d->fTheGammaData || p->fNumRegions < 1 || !p->fParametersPerRegion)
    return false;
  const int nmc = d->fTheMatCutData->fNumMatCutData;
  if (nmc < 1 || !d->fTheMatCutData->fMatCutData) return false;
  for (int i = 0; i < nmc; ++i) {
    const G4HepEmMCCData& c = d->fTheMatCutData->fMatCutData[i];
    if (c.fHepEmMatIndex < 0 || c.fHepEmMatIndex >= d->fTheMaterialData->fNumMaterialData) return false;
    if (c.fG4RegionIndex < 0 || c.fG4RegionIndex >= p->fNumRegions) return false;
  }
  return HepemTablesValid(d);
}

}  // namespace

// ---- the work buffer of the allocation-free step (hepem_step_noalloc) --------------------------
// One array per GPU, in the code object's device memory: the step's counters,
// particle-type queue, process codes, staged records and tile layout for up to
// HEPEM_NOALLOC_MAX_TRACKS slots (about 17.7 MiB), held until the process ends.
// In host code `scratch` names a host-side stand-in: only the address that
// hipGetSymbolAddress returns for the current device may reach a kernel.
// work layout (each part at 256-byte alignment): 64 counters | particle queue
// (n x u32) | process code per slot (n x u8) | staged record per slot
// (n x 64 bytes) | tile list (n x u16: each pending slot's offset within its
// tile, laid out by process) | tile info (kTileInfo words per tile of
// kClassifyTile slots)
namespace {
constexpr uint64_t kCountersBytes = 256;
static_assert(kCountersBytes == 4 * kCounterWords, "the counter words fill the counters' part of a work buffer");
constexpr uint64_t Al(uint64_t b) { return (b + 255) & ~255ull; }
// bytes of the work buffer of a step over m slots
constexpr uint64_t WorkBytes(uint64_t m) {
  return kCountersBytes + Al(4 * m) + Al(m) + Al(sizeof(Staged) * m) + Al(2 * m) +
         Al(4 * kTileInfo * ((m + kClassifyTile - 1) / kClassifyTile));
}
}  // namespace

namespace {
constexpr uint64_t kNoAllocScratchBytes = WorkBytes(HEPEM_NOALLOC_MAX_TRACKS);
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
// Device address of `scratch` on device `dev`, which must be the current
// device; looked up once per device and kept.
int ScratchAddress(int dev, void** ptr) {
  *ptr = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return HEPEM_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratchPtr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HEPEM_ERR_HIP;
    gScratchPtr[dev] = p;
  }
  *ptr = gScratchPtr[dev];
  return HEPEM_OK;
}
}  // namespace

struct hepem_ctx {
  int device = 0;
  int precision = HEPEM_FP64;
  void* image = nullptr;
  uint64_t imageBytes = 0;
  DevTables<double> t64;
  DevTables<float> t32;
  void* work = nullptr;       // internal work buffer for callers passing work == NULL
  uint64_t workBytes = 0;
  void* scratch = nullptr;    // device address of `scratch` on this context's device (hepem_step_noalloc)
  int scratchStatus = HEPEM_OK;
  int lookupStaging = 1;  // config-1 lookups read the tables from LDS where they fit
  int lookupGrid[2][2] = {{0, 0}, {0, 0}};  // [FP32 form][particle]: resident blocks of the staged kernel (0: not yet known)
  int processSort = HEPEM_SORT_AUTO;  // sort pending interactions by process (two stages): HEPEM_SORT_*
  // per couple: e-, e+, gamma production thresholds and the region's apply-cuts flag
  std::vector<double> cuts;
  std::vector<uint8_t> applyCuts;
  // per couple: G4HepEm material index and its density
  std::vector<int> coupleMat;
  std::vector<double> coupleDensity;
};

extern "C" const char* hepem_status_string(int s) {
  switch (s) {
    case HEPEM_OK: return "ok";
    case HEPEM_ERR_INVALID: return "invalid argument";
    case HEPEM_ERR_HIP: return "HIP runtime error";
    case HEPEM_ERR_CAPACITY: return "capacity exceeded";
    case HEPEM_ERR_ALLOC: return "allocation failed";
    case HEPEM_ERR_NO_DEVICE: return "no usable GPU device";
    case HEPEM_ERR_UNSUPPORTED: return "unsupported option";
  }
  return "unknown status";
}

extern "C" int hepem_ctx_create(const G4HepEmData* d, const G4HepEmParameters* p, int precision, hepem_ctx** out) {
  if (!out) return HEPEM_ERR_INVALID;
  *out = nullptr;
  if (!DataComplete(d, p) || (precision != HEPEM_FP64 && precision != HEPEM_FP32T)) return HEPEM_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return HEPEM_ERR_NO_DEVICE;
  hepem_ctx* c = new (std::nothrow) hepem_ctx();
  if (!c) return HEPEM_ERR_ALLOC;
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return HEPEM_ERR_HIP; }
  c->scratchStatus = ScratchAddress(c->device, &c->scratch);
  c->precision = precision;
  c->processSort = HEPEM_SORT_AUTO;
  {
    const G4HepEmMatCutData* mc = d->fTheMatCutData;
    c->cuts.resize(3 * (size_t)mc->fNumMatCutData);
    c->applyCuts.resize((size_t)mc->fNumMatCutData);
    for (int i = 0; i < mc->fNumMatCutData; ++i) {
      const G4HepEmMCCData& m = mc->fMatCutData[i];
      c->cuts[3 * (size_t)i] = m.fSecElProdCutE;
      c->cuts[3 * (size_t)i + 1] = m.fSecPosProdCutE;
      c->cuts[3 * (size_t)i + 2] = m.fSecGamProdCutE;
      const int r = m.fG4RegionIndex;
      c->applyCuts[(size_t)i] = (r >= 0 && r < p->fNumRegions && p->fParametersPerRegion[r].fIsApplyCuts) ? 1 : 0;
      c->coupleMat.push_back(m.fHepEmMatIndex);
      c->coupleDensity.push_back(d->fTheMaterialData->fMaterialData[m.fHepEmMatIndex].fDensity);
    }
  }
  Packer P;
  PackTables(P, c->t64, d, p);
  PackTables(P, c->t32, d, p);
  c->imageBytes = P.buf.size();
  if (hipMalloc(&c->image, c->imageBytes) != hipSuccess) { delete c; return HEPEM_ERR_ALLOC; }
  if (hipMemcpy(c->image, P.buf.data(), c->imageBytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(c->image);
    delete c;
    return HEPEM_ERR_HIP;
  }
  for (auto& f : P.fix) *f.first = static_cast<const uint8_t*>(c->image) + f.second;
  *out = c;
  return HEPEM_OK;
}

int HepemPrecision(const hepem_ctx* c) { return c->precision; }
const DevTables<double>& HepemTables64(const hepem_ctx* c) { return c->t64; }
const DevTables<float>& HepemTables32(const hepem_ctx* c) { return c->t32; }

int HepemCoupleMaterial(const hepem_ctx* c, int imc, int* imat, double* density) {
  if (!c || !imat || !density || imc < 0 || (size_t)imc >= c->coupleMat.size()) return HEPEM_ERR_INVALID;
  *imat = c->coupleMat[(size_t)imc];
  *density = c->coupleDensity[(size_t)imc];
  return HEPEM_OK;
}

int HepemCoupleCuts(const hepem_ctx* c, int imc, double cuts[3], int* apply) {
  if (!c || !cuts || !apply || imc < 0 || (size_t)imc >= c->applyCuts.size()) return HEPEM_ERR_INVALID;
  for (int k = 0; k < 3; ++k) cuts[k] = c->cuts[3 * (size_t)imc + k];
  *apply = c->applyCuts[(size_t)imc];
  return HEPEM_OK;
}

extern "C" int hepem_ctx_destroy(hepem_ctx* c) {
  if (!c) return HEPEM_ERR_INVALID;
  int rc = HEPEM_OK;
  if (c->image && hipFree(c->image) != hipSuccess) rc = HEPEM_ERR_HIP;
  if (c->work && hipFree(c->work) != hipSuccess) rc = HEPEM_ERR_HIP;
  delete c;
  return rc;
}

extern "C" int hepem_ctx_table_bytes(const hepem_ctx* c, uint64_t* bytes) {
  if (!c || !bytes) return HEPEM_ERR_INVALID;
  *bytes = c->imageBytes;
  return HEPEM_OK;
}

extern "C" int hepem_lookup_eloss_xsec(const hepem_ctx* c, int ptype, int32_t n, const int32_t* imc, const double* ekin,
                                       const double* lekin, double* out, void* stream) {
  if (!c || (ptype != HEPEM_ELECTRON && ptype != HEPEM_POSITRON) || n < 0) return HEPEM_ERR_INVALID;
  if (n == 0) return HEPEM_OK;
  if (!imc || !ekin || !lekin || !out) return HEPEM_ERR_INVALID;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = (n + kBlock - 1) / kBlock;
  const bool fp64 = c->precision == HEPEM_FP64;
  const size_t lds = fp64 ? LookupLdsBytes(c->t64.el[ptype], c->t64.nmc) : LookupLdsBytes(c->t32.el[ptype], c->t32.nmc);
  if (c->lookupStaging && lds <= kLookupLdsMax) {
    // blocks of kLookupPerThread queries per thread, fewer per thread when
    // that leaves less than one resident round of blocks; each block stages
    // the tables once
    int& resident = const_cast<hepem_ctx*>(c)->lookupGrid[fp64 ? 0 : 1][ptype];
    if (resident == 0) {
      int cus = 0, perCu = 0;
      HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
      if (fp64)
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_lookup_lds<double>, kBlock, lds));
      else
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_lookup_lds<float>, kBlock, lds));
      resident = cus * perCu > 0 ? cus * perCu : kLookupBlocks;
    }
    const int64_t perThread = std::max<int64_t>(1, std::min<int64_t>(kLookupPerThread, (int64_t)grid / resident));
    const int sgrid = (int)(((int64_t)n + kBlock * perThread - 1) / (kBlock * perThread));
    if (fp64)
      hipLaunchKernelGGL(k_lookup_lds<double>, dim3(sgrid), dim3(kBlock), lds, s, c->t64, ptype, n, imc, ekin, lekin, out);
    else
      hipLaunchKernelGGL(k_lookup_lds<float>, dim3(sgrid), dim3(kBlock), lds, s, c->t32, ptype, n, imc, ekin, lekin, out);
  } else if (fp64) {
    hipLaunchKernelGGL(k_lookup<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, ptype, n, imc, ekin, lekin, out);
  } else {
    hipLaunchKernelGGL(k_lookup<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, ptype, n, imc, ekin, lekin, out);
  }
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}

extern "C" int hepem_lookup_gamma_xsec(const hepem_ctx* c, int32_t n, const int32_t* imc, const double* ekin,
                                       const double* lekin, const double* urnd, double* out_xs, double* out_pe,
                                       int32_t* out_pid, void* stream) {
  if (!c || n < 0) return HEPEM_ERR_INVALID;
  if (n == 0) return HEPEM_OK;
  if (!imc || !ekin || !lekin || !urnd || !out_xs || !out_pe || !out_pid) return HEPEM_ERR_INVALID;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = (n + kBlock - 1) / kBlock;
  if (c->precision == HEPEM_FP64)
    hipLaunchKernelGGL(k_gamma_lookup<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, n, imc, ekin, lekin, urnd, out_xs,
                       out_pe, out_pid);
  else
    hipLaunchKernelGGL(k_gamma_lookup<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, n, imc, ekin, lekin, urnd, out_xs,
                       out_pe, out_pid);
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}

extern "C" int hepem_ctx_set_lookup_staging(hepem_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return HEPEM_ERR_INVALID;
  c->lookupStaging = on;
  return HEPEM_OK;
}

namespace {
// every array of a non-empty batch must be given; an empty batch may pass nulls
bool TracksOk(const hepem_tracks* t) {
  return t && t->n >= 0 &&
         (t->n == 0 || (t->ptype && t->id && t->imc && t->ekin && t->dir && t->safety && t->on_boundary && t->num_ia_left &&
                        t->msc_init_range && t->msc_dyn_range_factor && t->msc_tlimit_min && t->msc_first_step && t->rng_ctr &&
                        t->rng_gauss && t->rng_has_gauss));
}
bool ResOk(const hepem_step_out* r, int n) {
  return r && (n == 0 || (r->true_step && r->geom_step && r->edep && r->displacement && r->winner && r->stopped && r->nsec &&
                          r->sec_ptype && r->sec_ekin && r->sec_dir));
}
}  // namespace

// the parts of a work buffer, in the layout WorkBytes describes
namespace {
struct WorkLayout {
  uint32_t* counters;
  uint32_t* queue;
  uint8_t* code;
  Staged* st;
  uint16_t* tlist;   // the tile-local process layout: per slot, a tile offset
  uint32_t* tinfo;   // per classification tile: process offsets and first chunks
  WorkLayout(void* w, int32_t n) {
    uint8_t* p = static_cast<uint8_t*>(w);
    const uint64_t m = (uint64_t)(n > 0 ? n : 0);
    counters = reinterpret_cast<uint32_t*>(p);
    p += kCountersBytes;
    queue = reinterpret_cast<uint32_t*>(p);
    p += Al(4 * m);
    code = p;
    p += Al(m);
    st = reinterpret_cast<Staged*>(p);
    p += Al(sizeof(Staged) * m);
    tlist = reinterpret_cast<uint16_t*>(p);
    p += Al(2 * m);
    tinfo = reinterpret_cast<uint32_t*>(p);
  }
};
}  // namespace

extern "C" uint64_t hepem_step_work_bytes(int32_t n) { return WorkBytes((uint64_t)(n > 0 ? n : 0)); }

extern "C" int hepem_ctx_set_process_sort(hepem_ctx* c, int mode) {
  if (!c || (mode != HEPEM_SORT_OFF && mode != HEPEM_SORT_ON && mode != HEPEM_SORT_AUTO)) return HEPEM_ERR_INVALID;
  c->processSort = mode;
  return HEPEM_OK;
}

namespace {
// whether a step, batch plan or split Perform sorts its pending interactions
// by process; HEPEM_SORT_AUTO: the faster choice as measured, which is the
// sort at every size in both forms
bool SortsProcesses(const hepem_ctx* c) { return c->processSort != HEPEM_SORT_OFF; }
}  // namespace

namespace {
// The particle-type queue of a batch in `work`: over the whole batch (e-/e+
// from the front, gammas from the back; counters[0..1]), or, for the two-stage
// step, tile by tile (k_classify_local: each tile's e-/e+ first, then its
// gammas; the tile's e-/e+ count in its first tinfo word).
int PrepareQueue(const hepem_ctx* c, const hepem_tracks* in, void* work, void* stream, bool tileLocal) {
  if (!c || !TracksOk(in) || !work) return HEPEM_ERR_INVALID;
  const int n = in->n;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const WorkLayout w(work, n);
  const int grid = (n + kClassifyTile - 1) / kClassifyTile;
  if (tileLocal && n > 0) {
    // clears the counters itself
    hipLaunchKernelGGL(k_classify_local, dim3(grid), dim3(kClassifyBlock), 0, s, n, in->ptype, w.queue, w.tinfo, w.counters);
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kCounterWords), 0, s, w.counters, kCounterWords, 0u);
    if (n > 0)
      hipLaunchKernelGGL(k_classify, dim3(grid), dim3(kClassifyBlock), 0, s, n, in->ptype, w.queue, w.counters);
  }
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}
}  // namespace

extern "C" int hepem_step_prepare(const hepem_ctx* c, const hepem_tracks* in, void* work, void* stream) {
  if (!c || !TracksOk(in)) return HEPEM_ERR_INVALID;
  return PrepareQueue(c, in, work, stream, SortsProcesses(c));
}

extern "C" int hepem_step_run(const hepem_ctx* c, uint64_t seed, const hepem_tracks* in, hepem_tracks* out,
                              hepem_step_out* res, void* work, void* stream) {
  if (!c || !TracksOk(in) || !TracksOk(out) || !ResOk(res, in->n) || out->n != in->n || !work) return HEPEM_ERR_INVALID;
  const int n = in->n;
  if (n == 0) return HEPEM_OK;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const WorkLayout w(work, n);
  const int grid = (n + kBlock - 1) / kBlock;
  const bool fp64 = c->precision == HEPEM_FP64;
  if (!SortsProcesses(c)) {
    if (fp64)
      hipLaunchKernelGGL(k_step<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *in, *out, *res, w.queue, w.counters,
                         w.tinfo);
    else
      hipLaunchKernelGGL(k_step<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *in, *out, *res, w.queue, w.counters,
                         w.tinfo);
  } else {
    const int tiles = (n + kClassifyTile - 1) / kClassifyTile;
    const int igrid = InteractGrid(n);
    if (fp64) {
      hipLaunchKernelGGL(k_step_select_local<double>, dim3(LocalSelectGrid(n)), dim3(kBlock), 0, s, c->t64, seed, *in, *out, *res, w.queue,
                         w.counters, w.tinfo, w.code, w.st);
      hipLaunchKernelGGL(k_classify_tiles, dim3(tiles), dim3(kClassifyBlock), 0, s, n, w.code, w.tlist, w.tinfo);
      hipLaunchKernelGGL(k_interact_tiles<double>, dim3(igrid), dim3(kBlock), 0, s, c->t64, seed, *out, *res, w.code, w.st,
                         w.tlist, w.tinfo);
    } else {
      hipLaunchKernelGGL(k_step_select_local<float>, dim3(LocalSelectGrid(n)), dim3(kBlock), 0, s, c->t32, seed, *in, *out, *res, w.queue,
                         w.counters, w.tinfo, w.code, w.st);
      hipLaunchKernelGGL(k_classify_tiles, dim3(tiles), dim3(kClassifyBlock), 0, s, n, w.code, w.tlist, w.tinfo);
      hipLaunchKernelGGL(k_interact_tiles<float>, dim3(igrid), dim3(kBlock), 0, s, c->t32, seed, *out, *res, w.code, w.st,
                         w.tlist, w.tinfo);
    }
  }
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}

namespace {
// The context's own work buffer, grown to hold a queue of n slots.
int EnsureWork(const hepem_ctx* cc, int n, void*& work) {
  if (work) return HEPEM_OK;
  hepem_ctx* c = const_cast<hepem_ctx*>(cc);
  const uint64_t need = hepem_step_work_bytes(n);
  if (c->workBytes < need) {
    if (c->work) HIP_OK(hipFree(c->work));
    c->work = nullptr;
    c->workBytes = 0;
    if (hipMalloc(&c->work, need) != hipSuccess) return HEPEM_ERR_ALLOC;
    c->workBytes = need;
  }
  work = c->work;
  return HEPEM_OK;
}
}  // namespace

extern "C" int hepem_step(const hepem_ctx* c, uint64_t seed, const hepem_tracks* in, hepem_tracks* out, hepem_step_out* res,
                          void* work, void* stream) {
  if (!c || !TracksOk(in) || !TracksOk(out) || !ResOk(res, in->n) || out->n != in->n) return HEPEM_ERR_INVALID;
  const int n = in->n;
  if (n == 0) return HEPEM_OK;
  int rc = EnsureWork(c, n, work);
  if (rc != HEPEM_OK) return rc;
  rc = hepem_step_prepare(c, in, work, stream);
  if (rc != HEPEM_OK) return rc;
  return hepem_step_run(c, seed, in, out, res, work, stream);
}

extern "C" int hepem_step_noalloc(const hepem_ctx* c, uint64_t seed, const hepem_tracks* in, hepem_tracks* out,
                                  hepem_step_out* res, void* stream) {
  if (!c || !TracksOk(in) || !TracksOk(out) || !ResOk(res, in->n) || out->n != in->n) return HEPEM_ERR_INVALID;
  if (in->n > HEPEM_NOALLOC_MAX_TRACKS) return HEPEM_ERR_CAPACITY;
  if (in->n == 0) return HEPEM_OK;
  if (c->scratchStatus != HEPEM_OK) return c->scratchStatus;
  // the same kernels as hepem_step; prepare clears the counters (in the
  // particle classification, or with a fill kernel), so the scratch's previous
  // contents do not matter
  const int rc = hepem_step_prepare(c, in, c->scratch, stream);
  if (rc != HEPEM_OK) return rc;
  return hepem_step_run(c, seed, in, out, res, c->scratch, stream);
}

namespace {
bool HowFarOk(const hepem_howfar_state* h, int n) {
  return h && h->n == n &&
         (n == 0 || (h->geom_step && h->range && h->mfp && h->winner && h->msc_lambtr1 && h->msc_true_step && h->msc_zpath &&
                     h->msc_par1 && h->msc_par2 && h->msc_par3 && h->msc_active && h->msc_displace && h->pe_xsec));
}

// HowFar (perform == false) or Perform of a batch of slots of one KIND; a mixed
// batch is ordered by the particle-type queue in `work` (built here for HowFar).
template <int KIND>
int SplitCall(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs, hepem_step_out* res,
              void* work, void* stream, bool perform) {
  if (!c || !TracksOk(t) || !HowFarOk(hs, t ? t->n : -1) || (perform && !ResOk(res, t->n))) return HEPEM_ERR_INVALID;
  const int n = t->n;
  if (n == 0) return HEPEM_OK;
  int rc = EnsureWork(c, n, work);
  if (rc != HEPEM_OK) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const WorkLayout wl(work, n);
  uint32_t* counters = wl.counters;
  const uint32_t* queue = KIND == kMixed ? wl.queue : nullptr;
  if (!perform) {
    if (KIND == kMixed) {
      // the particle layout the coming Perform sorts best with, recorded in
      // the work buffer for the Perform kernels
      rc = PrepareQueue(c, t, work, stream, SortsProcesses(c));
      if (rc != HEPEM_OK) return rc;
    } else {
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counters, 64, 0u);
    }
  }
  // a mixed batch's kernels cover either layout of its queue
  const int grid = KIND == kMixed ? LocalSelectGrid(n) : (n + kBlock - 1) / kBlock;
  const bool fp64 = c->precision == HEPEM_FP64;
  if (!perform && fp64) {
    hipLaunchKernelGGL((k_howfar<double, KIND>), dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, queue, counters,
                       wl.tinfo);
  } else if (!perform) {
    hipLaunchKernelGGL((k_howfar<float, KIND>), dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, queue, counters,
                       wl.tinfo);
  } else if (SortsProcesses(c)) {
    // choice of the interaction, then the tile-local process sort and final
    // states (as hepem_step_run)
    const int tiles = (n + kClassifyTile - 1) / kClassifyTile;
    const int igrid = InteractGrid(n);
    if (fp64)
      hipLaunchKernelGGL((k_perform_select<double, KIND>), dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, *res, queue,
                         counters, wl.tinfo, wl.code, wl.st);
    else
      hipLaunchKernelGGL((k_perform_select<float, KIND>), dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, *res, queue,
                         counters, wl.tinfo, wl.code, wl.st);
    hipLaunchKernelGGL(k_classify_tiles, dim3(tiles), dim3(kClassifyBlock), 0, s, n, wl.code, wl.tlist, wl.tinfo);
    if (fp64)
      hipLaunchKernelGGL(k_interact_tiles<double>, dim3(igrid), dim3(kBlock), 0, s, c->t64, seed, *t, *res, wl.code, wl.st,
                         wl.tlist, wl.tinfo);
    else
      hipLaunchKernelGGL(k_interact_tiles<float>, dim3(igrid), dim3(kBlock), 0, s, c->t32, seed, *t, *res, wl.code, wl.st,
                         wl.tlist, wl.tinfo);
  } else if (fp64) {
    hipLaunchKernelGGL((k_perform<double, KIND>), dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, *res, queue, counters,
                       wl.tinfo);
  } else {
    hipLaunchKernelGGL((k_perform<float, KIND>), dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, *res, queue, counters,
                       wl.tinfo);
  }
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}
}  // namespace

extern "C" int hepem_howfar(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs, void* work,
                            void* stream) {
  return SplitCall<kMixed>(c, seed, t, hs, nullptr, work, stream, false);
}
extern "C" int hepem_perform(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                             hepem_step_out* res, void* work, void* stream) {
  return SplitCall<kMixed>(c, seed, t, hs, res, work, stream, true);
}
extern "C" int hepem_electron_howfar(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs, void* work,
                                     void* stream) {
  return SplitCall<kOnlyElectrons>(c, seed, t, hs, nullptr, work, stream, false);
}
extern "C" int hepem_electron_perform(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                                      hepem_step_out* res, void* work, void* stream) {
  return SplitCall<kOnlyElectrons>(c, seed, t, hs, res, work, stream, true);
}
extern "C" int hepem_gamma_howfar(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs, void* work,
                                  void* stream) {
  return SplitCall<kOnlyGammas>(c, seed, t, hs, nullptr, work, stream, false);
}
extern "C" int hepem_gamma_perform(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                                   hepem_step_out* res, void* work, void* stream) {
  return SplitCall<kOnlyGammas>(c, seed, t, hs, res, work, stream, true);
}

namespace {
bool SubstepOk(const hepem_substep_state* ss, int n) {
  return ss && ss->n == n &&
         (n == 0 || (ss->discrete_winner && ss->step_left && ss->true_path && ss->geom_path && ss->mean_eloss && ss->pre_ekin &&
                     ss->displacement && ss->active && ss->more && ss->stopped && ss->left_world && ss->nsub));
}

// the sub-step calls: 0 begin, 1 MSC HowFar, 2 along, 3 end
int SubstepCall(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs, hepem_substep_state* ss,
                hepem_step_out* res, void* work, void* stream, int phase) {
  if (!c || !TracksOk(t) || !HowFarOk(hs, t ? t->n : -1) || !SubstepOk(ss, t ? t->n : -1) ||
      (phase == 3 && !ResOk(res, t->n)))
    return HEPEM_ERR_INVALID;
  const int n = t->n;
  int rc = EnsureWork(c, n > 0 ? n : 1, work);
  if (rc != HEPEM_OK) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t* counters = static_cast<uint32_t*>(work);
  if (phase == 0) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counters, 64, 0u);
  if (phase == 2) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counters + 4, 1, 0u);
  if (n == 0) {
    HIP_OK(hipGetLastError());
    return HEPEM_OK;
  }
  const int grid = (n + kBlock - 1) / kBlock;
  const bool fp64 = c->precision == HEPEM_FP64;
  switch (phase) {
    case 0:
      if (fp64) hipLaunchKernelGGL(k_el_begin<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, *ss, counters);
      else hipLaunchKernelGGL(k_el_begin<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, *ss, counters);
      break;
    case 1:
      if (fp64) hipLaunchKernelGGL(k_el_substep_howfar<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, *ss);
      else hipLaunchKernelGGL(k_el_substep_howfar<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, *ss);
      break;
    case 2:
      if (fp64)
        hipLaunchKernelGGL(k_el_substep_along<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, *ss, counters);
      else
        hipLaunchKernelGGL(k_el_substep_along<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, *ss, counters);
      break;
    default:
      if (fp64) hipLaunchKernelGGL(k_el_end<double>, dim3(grid), dim3(kBlock), 0, s, c->t64, seed, *t, *hs, *ss, *res);
      else hipLaunchKernelGGL(k_el_end<float>, dim3(grid), dim3(kBlock), 0, s, c->t32, seed, *t, *hs, *ss, *res);
      break;
  }
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}
}  // namespace

extern "C" int hepem_electron_step_begin(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                                         hepem_substep_state* ss, void* work, void* stream) {
  return SubstepCall(c, seed, t, hs, ss, nullptr, work, stream, 0);
}
extern "C" int hepem_electron_substep_howfar(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                                             hepem_substep_state* ss, void* work, void* stream) {
  return SubstepCall(c, seed, t, hs, ss, nullptr, work, stream, 1);
}
extern "C" int hepem_electron_substep_along(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                                            hepem_substep_state* ss, void* work, void* stream) {
  return SubstepCall(c, seed, t, hs, ss, nullptr, work, stream, 2);
}
extern "C" int hepem_electron_step_end(const hepem_ctx* c, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                                       hepem_substep_state* ss, hepem_step_out* res, void* work, void* stream) {
  return SubstepCall(c, seed, t, hs, ss, res, work, stream, 3);
}
extern "C" int hepem_electron_substep_active(const void* work, uint32_t* nactive, void* stream) {
  if (!work || !nactive) return HEPEM_ERR_INVALID;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_OK(hipMemcpyAsync(nactive, static_cast<const uint32_t*>(work) + 4, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return HEPEM_OK;
}

extern "C" int hepem_step_rejected(const void* work, uint32_t* nrej, void* stream) {
  if (!work || !nrej) return HEPEM_ERR_INVALID;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_OK(hipMemcpyAsync(nrej, static_cast<const uint32_t*>(work) + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return HEPEM_OK;
}

// ---- batch plans ----------------------------------------------------------------------------------
struct hepem_batch_plan {
  const hepem_ctx* ctx = nullptr;
  int nbatch = 0;
  uint64_t total = 0;             // tracks over all batches
  uint32_t ntiles = 0;            // particle-classify tiles over all batches (kClassifyTile slots)
  uint32_t nptiles[kNumTiles] = {0, 0, 0, 0};  // process-sort tiles over all batches, per tile size
  uint32_t nchunks[kNumShapes] = {0, 0, 0};
  BatchDev* dDescs = nullptr;
  uint32_t* dQueue = nullptr;
  uint32_t* dCounters = nullptr;  // 4 per batch, then kQueueWords work counters
  // process sort of the pending interactions over all batches (as hepem_step)
  int sortMode = HEPEM_SORT_AUTO; // the context's mode at plan creation: what the tuner may choose from
  bool processSort = true;
  uint8_t* dCode = nullptr;       // process code per global slot
  Staged* dStaged = nullptr;      // staged record per global slot
  uint16_t* dTList = nullptr;     // tile-local process layout, batch b's slots from qoff_b
  uint32_t* dTInfo = nullptr;     // kTileInfo words per process-sort tile (tiles numbered across batches)
  uint32_t* dLocal = nullptr;     // e-/e+ count per particle-classify tile of the tile-local partition
  uint32_t* dTileBatch = nullptr; // batch of every tile, one array per tile size TileSlots(t), from tbOff[t]
  uint64_t tbOff[kNumTiles] = {0, 0, 0, 0};
  uint32_t* dChunkBatch = nullptr;  // batch of every step chunk, one array per block size ShapeBlock(s), from cbOff[s]
  uint64_t cbOff[kNumShapes] = {0, 0, 0};
  int shape = kNumShapes - 1;     // block size ShapeBlock(shape)
  int tile = 2;                   // process-sort tile size TileSlots(tile)
  bool persistent = false;        // step kernels as one resident round of blocks taking work from counters
  int grid = 0;                   // step-kernel blocks (nchunks[shape], or the resident round)
  int igrid = 0;                  // blocks of the persistent tile kernel (its resident round)
};

namespace {
// whether the plan's shape steps over the tile-local particle partition
// (sorted, dispatched, kBlock threads) or over each batch's particle queue
bool TileLocal(const hepem_batch_plan* p) { return p->processSort && !p->persistent && ShapeBlock(p->shape) == kBlock; }

// one resident round of blocks of `kernel` at `block` threads, a multiple of
// kXcds (the persistent tile kernel gives every XCD's counter its blocks)
template <typename K> int ResidentGrid(const hepem_ctx* c, K kernel, int block) {
  int cus = 0, perCu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, block, 0) != hipSuccess || cus * perCu <= 0)
    return 2048;
  return (cus * perCu + kXcds - 1) / kXcds * kXcds;
}

template <typename R, int S, int TI>
void LaunchStepBatched(const DevTables<R>& T, const hepem_batch_plan* p, uint64_t seed, hipStream_t s) {
  uint32_t* next = p->dCounters + 4 * p->nbatch;
  if (!p->processSort) {
    if (p->persistent)
      hipLaunchKernelGGL((k_step_batched_q<R, S>), dim3(p->grid), dim3(ShapeBlock(S)), 0, s, T, seed, p->dDescs, p->dChunkBatch + p->cbOff[S],
                         p->nchunks[S], p->dQueue, p->dCounters, next);
    else
      hipLaunchKernelGGL((k_step_batched<R, S>), dim3(p->grid), dim3(ShapeBlock(S)), 0, s, T, seed, p->dDescs, p->dChunkBatch + p->cbOff[S],
                         p->nchunks[S], p->dQueue, p->dCounters);
    return;
  }
  if (p->persistent)
    hipLaunchKernelGGL((k_step_select_batched_q<R, S>), dim3(p->grid), dim3(ShapeBlock(S)), 0, s, T, seed, p->dDescs,
                       p->dChunkBatch + p->cbOff[S], p->nchunks[S], p->dQueue, p->dCounters, p->dCode, p->dStaged, next);
  else if (TileLocal(p))
    hipLaunchKernelGGL((k_step_select_local_batched<R>), dim3(p->grid), dim3(kBlock), 0, s, T,
                       seed, p->dDescs, p->dTileBatch + p->tbOff[kParticleTiles], (int)p->ntiles, p->dQueue, p->dCounters,
                       p->dLocal, p->dCode, p->dStaged);
  else
    hipLaunchKernelGGL((k_step_select_batched<R, S>), dim3(p->grid), dim3(ShapeBlock(S)), 0, s, T, seed, p->dDescs, p->dChunkBatch + p->cbOff[S],
                       p->nchunks[S], p->dQueue, p->dCounters, p->dCode, p->dStaged);
  const uint32_t nt = p->nptiles[TI];
  hipLaunchKernelGGL(k_classify_tiles_batched<TI>, dim3(nt), dim3(kClassifyBlock), 0, s, p->dDescs, p->dTileBatch + p->tbOff[TI], p->dCode,
                     p->dTList, p->dTInfo);
  if (p->persistent) {
    hipLaunchKernelGGL((k_interact_tiles_batched_q<R, TI>), dim3(p->igrid), dim3(kBlock), 0, s, T, seed, p->dDescs,
                       p->dTileBatch + p->tbOff[TI], nt, p->dCode, p->dStaged, p->dTList, p->dTInfo, next + 1);
  } else {
    hipLaunchKernelGGL((k_interact_tiles_batched<R, TI>), dim3(InteractGridTiles((int)nt, InteractParts(TileItems(TI)))),
                       dim3(kBlock), 0, s, T, seed, p->dDescs, p->dTileBatch + p->tbOff[TI], nt, p->dCode, p->dStaged, p->dTList,
                       p->dTInfo);
  }
}
template <typename R, int S>
void LaunchStepBatched(const DevTables<R>& T, const hepem_batch_plan* p, uint64_t seed, hipStream_t s) {
  switch (p->tile) {
    case 0: LaunchStepBatched<R, S, 0>(T, p, seed, s); break;
    case 1: LaunchStepBatched<R, S, 1>(T, p, seed, s); break;
    case 3: LaunchStepBatched<R, S, 3>(T, p, seed, s); break;
    default: LaunchStepBatched<R, S, 2>(T, p, seed, s); break;
  }
}
template <typename R> void LaunchStepBatched(const DevTables<R>& T, const hepem_batch_plan* p, uint64_t seed, hipStream_t s) {
  switch (p->shape) {
    case 0: LaunchStepBatched<R, 0>(T, p, seed, s); break;
    case 1: LaunchStepBatched<R, 1>(T, p, seed, s); break;
    default: LaunchStepBatched<R, 2>(T, p, seed, s); break;
  }
}

// blocks of the step kernel (its first stage with the process sort) for the
// plan's shape: every chunk, the tile-local placement of every tile's chunks,
// or one resident round
template <typename R, int S> int StepGrid(const hepem_batch_plan* p) {
  if (!p->persistent && TileLocal(p)) return std::max(1, LocalSelectGridTiles((int)p->ntiles));
  if (!p->persistent) return (int)(p->nchunks[S] > 0 ? p->nchunks[S] : 1);
  return p->processSort ? ResidentGrid(p->ctx, k_step_select_batched_q<R, S>, ShapeBlock(S))
                        : ResidentGrid(p->ctx, k_step_batched_q<R, S>, ShapeBlock(S));
}
int StepGrid(const hepem_batch_plan* p) {
  const bool fp64 = p->ctx->precision == HEPEM_FP64;
  switch (p->shape) {
    case 0: return fp64 ? StepGrid<double, 0>(p) : StepGrid<float, 0>(p);
    case 1: return fp64 ? StepGrid<double, 1>(p) : StepGrid<float, 1>(p);
    default: return fp64 ? StepGrid<double, 2>(p) : StepGrid<float, 2>(p);
  }
}
template <typename R> int TileGrid(const hepem_batch_plan* p) {
  switch (p->tile) {
    case 0: return ResidentGrid(p->ctx, k_interact_tiles_batched_q<R, 0>, kBlock);
    case 1: return ResidentGrid(p->ctx, k_interact_tiles_batched_q<R, 1>, kBlock);
    case 3: return ResidentGrid(p->ctx, k_interact_tiles_batched_q<R, 3>, kBlock);
    default: return ResidentGrid(p->ctx, k_interact_tiles_batched_q<R, 2>, kBlock);
  }
}
// the grids of the plan's current shape
void SetGrids(hepem_batch_plan* p) {
  p->grid = StepGrid(p);
  p->igrid = !p->persistent ? 0 : p->ctx->precision == HEPEM_FP64 ? TileGrid<double>(p) : TileGrid<float>(p);
}

void FreePlan(hepem_batch_plan* p) {
  if (p->dDescs) (void)hipFree(p->dDescs);
  if (p->dQueue) (void)hipFree(p->dQueue);
  if (p->dCounters) (void)hipFree(p->dCounters);
  if (p->dCode) (void)hipFree(p->dCode);
  if (p->dStaged) (void)hipFree(p->dStaged);
  if (p->dTList) (void)hipFree(p->dTList);
  if (p->dTInfo) (void)hipFree(p->dTInfo);
  if (p->dLocal) (void)hipFree(p->dLocal);
  if (p->dTileBatch) (void)hipFree(p->dTileBatch);
  if (p->dChunkBatch) (void)hipFree(p->dChunkBatch);
  delete p;
}
}  // namespace

extern "C" int hepem_batch_plan_create(const hepem_ctx* c, const hepem_batch_desc* descs, int32_t nbatch,
                                       hepem_batch_plan** out) {
  if (!out) return HEPEM_ERR_INVALID;
  *out = nullptr;
  if (!c || nbatch < 1 || !descs) return HEPEM_ERR_INVALID;
  std::vector<BatchDev> h((size_t)nbatch + 1);
  uint64_t q = 0, tiles = 0, chunks[kNumShapes] = {0, 0, 0}, ptiles[kNumTiles] = {0, 0, 0, 0};
  for (int b = 0; b < nbatch; ++b) {
    const hepem_batch_desc& d = descs[b];
    if (!TracksOk(&d.in) || !TracksOk(&d.out) || !ResOk(&d.res, d.in.n) || d.out.n != d.in.n) return HEPEM_ERR_INVALID;
    BatchDev& x = h[b];
    x.in = d.in;
    x.out = d.out;
    x.res = d.res;
    x.qoff = (uint32_t)q;
    x.tile0 = (uint32_t)tiles;
    for (int t = 0; t < kNumTiles; ++t) x.ptile0[t] = (uint32_t)ptiles[t];
    for (int s = 0; s < kNumShapes; ++s) x.chunk0[s] = (uint32_t)chunks[s];
    const uint64_t n = (uint64_t)d.in.n;
    q += n;
    tiles += (n + kClassifyTile - 1) / kClassifyTile;
    for (int t = 0; t < kNumTiles; ++t) ptiles[t] += (n + TileSlots(t) - 1) / TileSlots(t);
    for (int s = 0; s < kNumShapes; ++s) chunks[s] += (n + ShapeBlock(s) - 1) / ShapeBlock(s);
  }
  if (q >= (1ull << 31)) return HEPEM_ERR_CAPACITY;
  // the batch of every tile, per tile size, and of every step chunk, per
  // block size: the kernels look it up instead of searching the descriptors'
  // prefix sums
  std::vector<uint32_t> tileBatch, chunkBatch;
  uint64_t tbOff[kNumTiles], cbOff[kNumShapes];
  for (int t = 0; t < kNumTiles; ++t) {
    tbOff[t] = tileBatch.size();
    for (int b = 0; b < nbatch; ++b) {
      const uint64_t k = (b + 1 < nbatch ? (uint64_t)h[b + 1].ptile0[t] : ptiles[t]) - h[b].ptile0[t];
      tileBatch.insert(tileBatch.end(), (size_t)k, (uint32_t)b);
    }
  }
  for (int sh = 0; sh < kNumShapes; ++sh) {
    cbOff[sh] = chunkBatch.size();
    for (int b = 0; b < nbatch; ++b) {
      const uint64_t k = (b + 1 < nbatch ? (uint64_t)h[b + 1].chunk0[sh] : chunks[sh]) - h[b].chunk0[sh];
      chunkBatch.insert(chunkBatch.end(), (size_t)k, (uint32_t)b);
    }
  }
  BatchDev& last = h[nbatch];
  std::memset(&last, 0, sizeof(last));
  last.qoff = (uint32_t)q;
  last.tile0 = (uint32_t)tiles;
  for (int t = 0; t < kNumTiles; ++t) last.ptile0[t] = (uint32_t)ptiles[t];
  for (int s = 0; s < kNumShapes; ++s) last.chunk0[s] = (uint32_t)chunks[s];

  hepem_batch_plan* p = new (std::nothrow) hepem_batch_plan();
  if (!p) return HEPEM_ERR_ALLOC;
  p->ctx = c;
  p->sortMode = c->processSort;
  p->processSort = SortsProcesses(c);
  p->nbatch = nbatch;
  p->total = q;
  p->ntiles = (uint32_t)tiles;
  for (int t = 0; t < kNumTiles; ++t) p->nptiles[t] = (uint32_t)ptiles[t];
  for (int s = 0; s < kNumShapes; ++s) p->nchunks[s] = (uint32_t)chunks[s];
  SetGrids(p);
  // the smallest tile size has the most tiles: its info words fit every size
  if (hipMalloc(&p->dDescs, h.size() * sizeof(BatchDev)) != hipSuccess ||
      hipMalloc(&p->dQueue, (q > 0 ? q : 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&p->dCounters, ((size_t)4 * nbatch + kQueueWords) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&p->dCode, (q > 0 ? q : 1)) != hipSuccess || hipMalloc(&p->dStaged, (q > 0 ? q : 1) * sizeof(Staged)) != hipSuccess ||
      hipMalloc(&p->dTList, (q > 0 ? q : 1) * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&p->dTInfo, (size_t)kTileInfo * (ptiles[0] > 0 ? ptiles[0] : 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&p->dLocal, (tiles > 0 ? tiles : 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&p->dTileBatch, std::max<size_t>(1, tileBatch.size()) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&p->dChunkBatch, std::max<size_t>(1, chunkBatch.size()) * sizeof(uint32_t)) != hipSuccess) {
    FreePlan(p);
    return HEPEM_ERR_ALLOC;
  }
  for (int t = 0; t < kNumTiles; ++t) p->tbOff[t] = tbOff[t];
  for (int sh = 0; sh < kNumShapes; ++sh) p->cbOff[sh] = cbOff[sh];
  if (hipMemcpy(p->dDescs, h.data(), h.size() * sizeof(BatchDev), hipMemcpyHostToDevice) != hipSuccess ||
      (!tileBatch.empty() && hipMemcpy(p->dTileBatch, tileBatch.data(), tileBatch.size() * sizeof(uint32_t),
                                       hipMemcpyHostToDevice) != hipSuccess) ||
      (!chunkBatch.empty() && hipMemcpy(p->dChunkBatch, chunkBatch.data(), chunkBatch.size() * sizeof(uint32_t),
                                        hipMemcpyHostToDevice) != hipSuccess)) {
    FreePlan(p);
    return HEPEM_ERR_HIP;
  }
  *out = p;
  return HEPEM_OK;
}

extern "C" int hepem_batch_plan_destroy(hepem_batch_plan* p) {
  if (!p) return HEPEM_ERR_INVALID;
  FreePlan(p);
  return HEPEM_OK;
}

extern "C" int hepem_batch_plan_run(const hepem_batch_plan* p, uint64_t seed, void* stream) {
  if (!p || !p->ctx) return HEPEM_ERR_INVALID;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int nc = 4 * p->nbatch + kQueueWords;
  hipLaunchKernelGGL(k_fill_u32, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p->dCounters, nc, 0u);
  if (p->total > 0) {
    if (TileLocal(p))
      hipLaunchKernelGGL(k_classify_local_batched, dim3(p->ntiles), dim3(kClassifyBlock), 0, s, p->dDescs,
                         p->dTileBatch + p->tbOff[kParticleTiles], p->dQueue, p->dLocal);
    else
      hipLaunchKernelGGL(k_classify_batched, dim3(p->ntiles), dim3(kClassifyBlock), 0, s, p->dDescs,
                         p->dTileBatch + p->tbOff[kParticleTiles], p->dQueue, p->dCounters);
    if (p->ctx->precision == HEPEM_FP64) LaunchStepBatched(p->ctx->t64, p, seed, s);
    else LaunchStepBatched(p->ctx->t32, p, seed, s);
  }
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}

namespace {
struct PlanShape {
  int shape, tile;
  bool sort, persistent;
};
void ApplyShape(hepem_batch_plan* p, const PlanShape& x) {
  p->shape = x.shape;
  p->tile = x.tile;
  p->processSort = x.sort;
  p->persistent = x.persistent;
  SetGrids(p);
}
}  // namespace

extern "C" int hepem_batch_plan_tune(hepem_batch_plan* p, uint64_t seed, void* stream) {
  if (!p || !p->ctx) return HEPEM_ERR_INVALID;
  if (p->total == 0) return HEPEM_OK;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // the shapes the plan may take: the sort modes its context allows, every
  // block size, every tile size when sorting, dispatched or persistent
  std::vector<PlanShape> cand;
  for (int so = 0; so < 2; ++so) {
    const bool sort = so == 1;
    if ((p->sortMode == HEPEM_SORT_OFF && sort) || (p->sortMode == HEPEM_SORT_ON && !sort)) continue;
    for (int sh = 0; sh < kNumShapes; ++sh)
      for (int t = 0; t < (sort ? kNumTiles : 1); ++t)
        for (int pe = 0; pe < 2; ++pe) cand.push_back({sh, sort ? t : 2, sort, pe == 1});
  }
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return HEPEM_ERR_HIP; }
  const PlanShape start{p->shape, p->tile, p->processSort, p->persistent};
  PlanShape best = start;
  float bestMs = 1e30f;
  int rc = HEPEM_OK;
  for (const PlanShape& x : cand) {
    ApplyShape(p, x);
    rc = hepem_batch_plan_run(p, seed, stream);  // warm-up
    if (rc != HEPEM_OK) break;
    if (hipEventRecord(e0, s) != hipSuccess) { rc = HEPEM_ERR_HIP; break; }
    for (int r = 0; r < 3 && rc == HEPEM_OK; ++r) rc = hepem_batch_plan_run(p, seed, stream);
    if (rc != HEPEM_OK || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = rc != HEPEM_OK ? rc : HEPEM_ERR_HIP;
      break;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = HEPEM_ERR_HIP; break; }
    if (ms < bestMs) {
      bestMs = ms;
      best = x;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ApplyShape(p, rc == HEPEM_OK ? best : start);
  return rc;
}

extern "C" int hepem_batch_plan_shape(const hepem_batch_plan* p, int32_t* block, int32_t* grid) {
  if (!p || !block || !grid) return HEPEM_ERR_INVALID;
  *block = ShapeBlock(p->shape);
  *grid = p->grid;
  return HEPEM_OK;
}

extern "C" int hepem_batch_plan_get_shape(const hepem_batch_plan* p, hepem_batch_shape* out) {
  if (!p || !out) return HEPEM_ERR_INVALID;
  out->block = ShapeBlock(p->shape);
  out->sort_tile = p->processSort ? TileSlots(p->tile) : 0;
  out->persistent = p->persistent ? 1 : 0;
  out->grid = p->grid;
  return HEPEM_OK;
}

extern "C" int hepem_batch_plan_set_shape(hepem_batch_plan* p, const hepem_batch_shape* in) {
  if (!p || !p->ctx || !in || (in->persistent != 0 && in->persistent != 1)) return HEPEM_ERR_INVALID;
  PlanShape x{-1, 2, in->sort_tile != 0, in->persistent == 1};
  for (int sh = 0; sh < kNumShapes; ++sh)
    if (in->block == ShapeBlock(sh)) x.shape = sh;
  if (x.shape < 0) return HEPEM_ERR_INVALID;
  if (x.sort) {
    x.tile = -1;
    for (int t = 0; t < kNumTiles; ++t)
      if (in->sort_tile == TileSlots(t)) x.tile = t;
    if (x.tile < 0) return HEPEM_ERR_INVALID;
  }
  ApplyShape(p, x);
  return HEPEM_OK;
}

extern "C" int hepem_batch_plan_rejected(const hepem_batch_plan* p, int32_t b, uint32_t* nrej, void* stream) {
  if (!p || !nrej || b < 0 || b >= p->nbatch) return HEPEM_ERR_INVALID;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_OK(hipMemcpyAsync(nrej, p->dCounters + 4 * b + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return HEPEM_OK;
}

extern "C" int hepem_step_batched(const hepem_ctx* c, uint64_t seed, const hepem_batch_desc* descs, int32_t nbatch,
                                  void* stream) {
  if (!c || nbatch < 0 || (nbatch > 0 && !descs)) return HEPEM_ERR_INVALID;
  if (nbatch == 0) return HEPEM_OK;
  hepem_batch_plan* p = nullptr;
  int rc = hepem_batch_plan_create(c, descs, nbatch, &p);
  if (rc != HEPEM_OK) return rc;
  rc = hepem_batch_plan_run(p, seed, stream);
  if (rc == HEPEM_OK && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess) rc = HEPEM_ERR_HIP;
  FreePlan(p);
  return rc;
}

extern "C" int hepem_tracks_reset_state(const hepem_tracks* t, void* stream) {
  if (!TracksOk(t)) return HEPEM_ERR_INVALID;
  if (t->n == 0) return HEPEM_OK;
  const int grid = (t->n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_reset_state, dim3(grid), dim3(kBlock), 0, static_cast<hipStream_t>(stream), *t);
  HIP_OK(hipGetLastError());
  return HEPEM_OK;
}
