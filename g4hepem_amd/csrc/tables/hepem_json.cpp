// hepem_json.cpp -- G4HepEmData / G4HepEmParameters to and from JSON text in
// the reference's schema (include/g4hepem_amd_tables.h).
//
// The schema -- key names, nesting, which counts are implied by array sizes,
// empty arrays written as null, only elements with Z > 0 persisted -- is the
// one of the reference's serializer (G4HepEm/G4HepEmDataJsonIO/src/
// G4HepEmDataJsonIOImpl.hh:121-979; entry points G4HepEmDataJsonIO.cc:8-61),
// so documents written by G4HepEmDataToJson / G4HepEmStateToJson load here and
// the other way round. The parser is a small streaming reader that fills the
// plain-C structs directly (no document tree): numbers through strtod (exactly
// rounded), doubles written with 17 significant digits, so a write/read round
// trip reproduces every value bit for bit. Memory is allocated with calloc,
// as the offline builder does, so hepem_tables_free releases it.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "g4hepem_amd.h"
#include "g4hepem_amd_tables.h"
#include "hepem_validate.hh"

namespace {

// ---- reading ----------------------------------------------------------------------------------

struct BadJson {};  // thrown on any syntax, type or schema error; mapped to HEPEM_ERR_INVALID

class In {
 public:
  In(const char* s, size_t n) : p_(s), e_(s + n) {}

  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
  }
  bool peek(char c) {
    ws();
    return p_ < e_ && *p_ == c;
  }
  void expect(char c) {
    if (!peek(c)) throw BadJson{};
    ++p_;
  }
  bool literal(const char* w) {
    ws();
    const size_t n = std::strlen(w);
    if ((size_t)(e_ - p_) >= n && std::memcmp(p_, w, n) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  bool atEnd() {
    ws();
    return p_ == e_;
  }

  std::string str() {
    expect('"');
    std::string out;
    while (p_ < e_ && *p_ != '"') {
      if (*p_ == '\\') {
        ++p_;
        if (p_ >= e_) throw BadJson{};
        switch (*p_) {
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          case '/': out += '/'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'n': out += '\n'; break;
          case 'r': out += '\r'; break;
          case 't': out += '\t'; break;
          case 'u':  // keys of the schema are plain ASCII: keep the escape as text
            if (e_ - p_ < 5) throw BadJson{};
            out.append(p_ - 1, 6);
            p_ += 4;
            break;
          default: throw BadJson{};
        }
        ++p_;
      } else {
        out += *p_++;
      }
    }
    expect('"');
    return out;
  }

  double num() {
    ws();
    if (p_ >= e_) throw BadJson{};
    // JSON number grammar, then strtod on a bounded copy
    const char* q = p_;
    if (q < e_ && *q == '-') ++q;
    if (q >= e_ || !(*q >= '0' && *q <= '9')) throw BadJson{};
    while (q < e_ && ((*q >= '0' && *q <= '9') || *q == '.' || *q == 'e' || *q == 'E' || *q == '+' || *q == '-')) ++q;
    const size_t n = (size_t)(q - p_);
    if (n > 64) throw BadJson{};
    char buf[72];
    std::memcpy(buf, p_, n);
    buf[n] = '\0';
    char* endp = nullptr;
    errno = 0;
    const double v = std::strtod(buf, &endp);
    if (endp != buf + n || !std::isfinite(v)) throw BadJson{};
    p_ = q;
    return v;
  }
  int inum() {
    const double v = num();
    if (!(v >= -2147483648.0 && v <= 2147483647.0) || v != std::floor(v)) throw BadJson{};
    return (int)v;
  }
  bool boolean() {
    if (literal("true")) return true;
    if (literal("false")) return false;
    throw BadJson{};
  }

  // null or an array of numbers
  template <typename T> std::vector<T> arr() {
    std::vector<T> v;
    if (literal("null")) return v;
    expect('[');
    if (peek(']')) {
      ++p_;
      return v;
    }
    while (true) {
      if constexpr (std::is_same<T, int>::value) v.push_back(inum());
      else v.push_back(num());
      if (peek(',')) {
        ++p_;
        continue;
      }
      expect(']');
      return v;
    }
  }
  // a fixed-size array of numbers
  template <typename T, size_t N> void fixed(T (&out)[N]) {
    const std::vector<T> v = arr<T>();
    if (v.size() != N) throw BadJson{};
    for (size_t i = 0; i < N; ++i) out[i] = v[i];
  }

  // Object: calls on(key) for every member; `on` consumes the value and
  // returns false for a key it does not know (that value is skipped).
  template <typename F> void object(F on) {
    expect('{');
    if (peek('}')) {
      ++p_;
      return;
    }
    while (true) {
      const std::string k = str();
      expect(':');
      if (!on(k)) skip();
      if (peek(',')) {
        ++p_;
        continue;
      }
      expect('}');
      return;
    }
  }
  // Array of objects (or null): calls each() once per element.
  template <typename F> void objects(F each) {
    if (literal("null")) return;
    expect('[');
    if (peek(']')) {
      ++p_;
      return;
    }
    while (true) {
      each();
      if (peek(',')) {
        ++p_;
        continue;
      }
      expect(']');
      return;
    }
  }
  bool null() { return literal("null"); }

  void skip(int depth = 0) {
    if (depth > 64) throw BadJson{};
    ws();
    if (p_ >= e_) throw BadJson{};
    if (*p_ == '"') {
      str();
    } else if (*p_ == '{') {
      object([&](const std::string&) {
        skip(depth + 1);
        return true;
      });
    } else if (*p_ == '[') {
      ++p_;
      if (peek(']')) {
        ++p_;
        return;
      }
      while (true) {
        skip(depth + 1);
        if (peek(',')) {
          ++p_;
          continue;
        }
        expect(']');
        return;
      }
    } else if (!literal("true") && !literal("false") && !literal("null")) {
      num();
    }
  }

 private:
  const char* p_;
  const char* e_;
};

// Required-key bookkeeping of one object.
class Keys {
 public:
  explicit Keys(std::initializer_list<const char*> names) : names_(names), seen_(names.size(), false) {}
  // marks `k` seen; true if it is one of the names
  bool mark(const std::string& k) {
    for (size_t i = 0; i < names_.size(); ++i)
      if (k == names_[i]) {
        if (seen_[i]) throw BadJson{};  // duplicate member
        seen_[i] = true;
        return true;
      }
    return false;
  }
  void require() const {
    for (bool s : seen_)
      if (!s) throw BadJson{};
  }

 private:
  std::vector<const char*> names_;
  std::vector<bool> seen_;
};

template <typename T> T* Copy(const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* p = static_cast<T*>(std::calloc(v.size(), sizeof(T)));
  if (!p) throw BadJson{};
  std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}
template <typename T> T* Zeroed(size_t n) {
  T* p = static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
  if (!p) throw BadJson{};
  return p;
}
int Count(size_t n) {
  if (n > 0x7FFFFFFFu) throw BadJson{};
  return (int)n;
}
void Check(bool ok) {
  if (!ok) throw BadJson{};
}

// G4HepEmDataJsonIOImpl.hh:121-160 (region) and :165-215 (parameters)
G4HepEmParameters* ReadParameters(In& in) {
  if (in.null()) return nullptr;
  G4HepEmParameters* p = Zeroed<G4HepEmParameters>(1);
  std::vector<G4HepEmRegionParmeters> regions;
  Keys ks{"fElectronTrackingCut", "fGammaTrackingCut", "fMinLossTableEnergy", "fMaxLossTableEnergy", "fNumLossTableBins",
          "fElectronBremModelLim", "fIsMSCPositronCor", "fIsMSCDisplacement", "fNumRegions", "fParametersPerRegion"};
  try {
    in.object([&](const std::string& k) {
      if (!ks.mark(k)) return false;
      if (k == "fElectronTrackingCut") p->fElectronTrackingCut = in.num();
      else if (k == "fGammaTrackingCut") p->fGammaTrackingCut = in.num();
      else if (k == "fMinLossTableEnergy") p->fMinLossTableEnergy = in.num();
      else if (k == "fMaxLossTableEnergy") p->fMaxLossTableEnergy = in.num();
      else if (k == "fNumLossTableBins") p->fNumLossTableBins = in.inum();
      else if (k == "fElectronBremModelLim") p->fElectronBremModelLim = in.num();
      else if (k == "fIsMSCPositronCor") p->fIsMSCPositronCor = in.boolean();
      else if (k == "fIsMSCDisplacement") p->fIsMSCDisplacement = in.boolean();
      else if (k == "fNumRegions") p->fNumRegions = in.inum();
      else
        in.objects([&] {
          G4HepEmRegionParmeters r{};
          Keys rk{"fFinalRange", "fDRoverRange", "fLinELossLimit", "fMSCRangeFactor", "fMSCSafetyFactor",
                  "fIsMSCMinimalStepLimit", "fIsELossFluctuation", "fIsMultipleStepsInMSCTrans", "fIsApplyCuts"};
          in.object([&](const std::string& q) {
            if (!rk.mark(q)) return false;
            if (q == "fFinalRange") r.fFinalRange = in.num();
            else if (q == "fDRoverRange") r.fDRoverRange = in.num();
            else if (q == "fLinELossLimit") r.fLinELossLimit = in.num();
            else if (q == "fMSCRangeFactor") r.fMSCRangeFactor = in.num();
            else if (q == "fMSCSafetyFactor") r.fMSCSafetyFactor = in.num();
            else if (q == "fIsMSCMinimalStepLimit") r.fIsMSCMinimalStepLimit = in.boolean();
            else if (q == "fIsELossFluctuation") r.fIsELossFluctuation = in.boolean();
            else if (q == "fIsMultipleStepsInMSCTrans") r.fIsMultipleStepsInMSCTrans = in.boolean();
            else r.fIsApplyCuts = in.boolean();
            return true;
          });
          rk.require();
          regions.push_back(r);
        });
      return true;
    });
    ks.require();
    Check(p->fNumRegions == Count(regions.size()) && p->fNumRegions > 0);
    p->fParametersPerRegion = Copy(regions);
  } catch (...) {
    std::free(p->fParametersPerRegion);
    std::free(p);
    throw;
  }
  return p;
}

// :268-303 (G4HepEmMatCutData); the per-couple records :322-341
void ReadMatCut(In& in, G4HepEmMatCutData*& out) {
  if (in.null()) return;
  out = Zeroed<G4HepEmMatCutData>(1);
  std::vector<int> idx;
  std::vector<G4HepEmMCCData> mcc;
  Keys ks{"fNumG4MatCuts", "fNumMatCutData", "fG4MCIndexToHepEmMCIndex", "fMatCutData"};
  in.object([&](const std::string& k) {
    if (!ks.mark(k)) return false;
    if (k == "fNumG4MatCuts") out->fNumG4MatCuts = in.inum();
    else if (k == "fNumMatCutData") out->fNumMatCutData = in.inum();
    else if (k == "fG4MCIndexToHepEmMCIndex") idx = in.arr<int>();
    else
      in.objects([&] {
        G4HepEmMCCData d{};
        Keys mk{"fSecElProdCutE", "fSecPosProdCutE", "fSecGamProdCutE", "fLogSecGamCutE", "fHepEmMatIndex", "fG4MatCutIndex",
                "fG4RegionIndex"};
        in.object([&](const std::string& q) {
          if (!mk.mark(q)) return false;
          if (q == "fSecElProdCutE") d.fSecElProdCutE = in.num();
          else if (q == "fSecPosProdCutE") d.fSecPosProdCutE = in.num();
          else if (q == "fSecGamProdCutE") d.fSecGamProdCutE = in.num();
          else if (q == "fLogSecGamCutE") d.fLogSecGamCutE = in.num();
          else if (q == "fHepEmMatIndex") d.fHepEmMatIndex = in.inum();
          else if (q == "fG4MatCutIndex") d.fG4MatCutIndex = in.inum();
          else d.fG4RegionIndex = in.inum();
          return true;
        });
        mk.require();
        mcc.push_back(d);
      });
    return true;
  });
  ks.require();
  Check(out->fNumG4MatCuts == Count(idx.size()) && out->fNumMatCutData == Count(mcc.size()));
  out->fG4MCIndexToHepEmMCIndex = Copy(idx);
  out->fMatCutData = Copy(mcc);
}

// :369-396 (one material) and :421-459 (the set)
void ReadMaterials(In& in, G4HepEmMaterialData*& out) {
  if (in.null()) return;
  out = Zeroed<G4HepEmMaterialData>(1);  // counts stay 0 until the arrays are attached
  std::vector<int> idx;
  std::vector<G4HepEmMatData> mats;
  int numG4 = 0, numData = 0;
  Keys ks{"fNumG4Material", "fNumMaterialData", "fG4MatIndexToHepEmMatIndex", "fMaterialData"};
  try {
    in.object([&](const std::string& k) {
      if (!ks.mark(k)) return false;
      if (k == "fNumG4Material") numG4 = in.inum();
      else if (k == "fNumMaterialData") numData = in.inum();
      else if (k == "fG4MatIndexToHepEmMatIndex") idx = in.arr<int>();
      else
        in.objects([&] {
          G4HepEmMatData m{};
          std::vector<int> el;
          std::vector<double> atoms, se, sc;
          Keys mk{"fG4MatIndex", "fElementVect", "fNumOfAtomsPerVolumeVect", "fDensity", "fDensityCorfactor",
                  "fElectronDensity", "fRadiationLength", "fMeanExEnergy", "fSandiaEnergies", "fSandiaCoefficients",
                  "fZeff", "fZeff23", "fZeffSqrt", "fUMSCPar", "fUMSCStepMinPars", "fUMSCTailCoeff", "fUMSCThetaCoeff"};
          in.object([&](const std::string& q) {
            if (!mk.mark(q)) return false;
            if (q == "fG4MatIndex") m.fG4MatIndex = in.inum();
            else if (q == "fElementVect") el = in.arr<int>();
            else if (q == "fNumOfAtomsPerVolumeVect") atoms = in.arr<double>();
            else if (q == "fDensity") m.fDensity = in.num();
            else if (q == "fDensityCorfactor") m.fDensityCorFactor = in.num();
            else if (q == "fElectronDensity") m.fElectronDensity = in.num();
            else if (q == "fRadiationLength") m.fRadiationLength = in.num();
            else if (q == "fMeanExEnergy") m.fMeanExEnergy = in.num();
            else if (q == "fSandiaEnergies") se = in.arr<double>();
            else if (q == "fSandiaCoefficients") sc = in.arr<double>();
            else if (q == "fZeff") m.fZeff = in.num();
            else if (q == "fZeff23") m.fZeff23 = in.num();
            else if (q == "fZeffSqrt") m.fZeffSqrt = in.num();
            else if (q == "fUMSCPar") m.fUMSCPar = in.num();
            else if (q == "fUMSCStepMinPars") in.fixed(m.fUMSCStepMinPars);
            else if (q == "fUMSCTailCoeff") in.fixed(m.fUMSCTailCoeff);
            else in.fixed(m.fUMSCThetaCoeff);
            return true;
          });
          mk.require();
          Check(atoms.size() == el.size() && sc.size() == 4 * se.size());
          m.fNumOfElement = Count(el.size());
          m.fNumOfSandiaIntervals = Count(se.size());
          m.fElementVect = Copy(el);
          m.fNumOfAtomsPerVolumeVect = Copy(atoms);
          m.fSandiaEnergies = Copy(se);
          m.fSandiaCoefficients = Copy(sc);
          mats.push_back(m);
        });
      return true;
    });
    ks.require();
    Check(numG4 == Count(idx.size()) && numData == Count(mats.size()));
    out->fG4MatIndexToHepEmMatIndex = Copy(idx);
    out->fMaterialData = Copy(mats);
    out->fNumG4Material = numG4;
    out->fNumMaterialData = numData;
  } catch (...) {
    for (G4HepEmMatData& m : mats) {
      std::free(m.fElementVect);
      std::free(m.fNumOfAtomsPerVolumeVect);
      std::free(m.fSandiaEnergies);
      std::free(m.fSandiaCoefficients);
    }
    std::free(out->fG4MatIndexToHepEmMatIndex);
    std::free(out->fMaterialData);
    out->fG4MatIndexToHepEmMatIndex = nullptr;
    out->fMaterialData = nullptr;
    out->fNumG4Material = out->fNumMaterialData = 0;
    throw;
  }
}

// :223-263 (one element) and :268-310 (the set, indexed by Z; absent Z keep fZet = -1)
void ReadElements(In& in, G4HepEmElementData*& out) {
  if (in.null()) return;
  out = Zeroed<G4HepEmElementData>(1);
  out->fMaxZet = 120;
  out->fElementData = Zeroed<G4HepEmElemData>(121);
  for (int z = 0; z < 121; ++z) out->fElementData[z].fZet = -1.0;
  in.objects([&] {
    G4HepEmElemData e{};
    std::vector<double> se, sc;
    Keys ek{"fZet", "fZet13", "fZet23", "fCoulomb", "fLogZ", "fZFactor1", "fDeltaMaxLow", "fDeltaMaxHigh", "fILVarS1",
            "fILVarS1Cond", "fSandiaEnergies", "fSandiaCoefficients", "fKShellBindingEnergy"};
    in.object([&](const std::string& q) {
      if (!ek.mark(q)) return false;
      if (q == "fZet") e.fZet = in.num();
      else if (q == "fZet13") e.fZet13 = in.num();
      else if (q == "fZet23") e.fZet23 = in.num();
      else if (q == "fCoulomb") e.fCoulomb = in.num();
      else if (q == "fLogZ") e.fLogZ = in.num();
      else if (q == "fZFactor1") e.fZFactor1 = in.num();
      else if (q == "fDeltaMaxLow") e.fDeltaMaxLow = in.num();
      else if (q == "fDeltaMaxHigh") e.fDeltaMaxHigh = in.num();
      else if (q == "fILVarS1") e.fILVarS1 = in.num();
      else if (q == "fILVarS1Cond") e.fILVarS1Cond = in.num();
      else if (q == "fSandiaEnergies") se = in.arr<double>();
      else if (q == "fSandiaCoefficients") sc = in.arr<double>();
      else e.fKShellBindingEnergy = in.num();
      return true;
    });
    ek.require();
    const int z = (int)e.fZet;
    Check(e.fZet == (double)z && z >= 1 && z <= 120 && sc.size() == 4 * se.size());
    G4HepEmElemData& slot = out->fElementData[z];
    Check(slot.fZet < 0.0);  // each Z once
    e.fNumOfSandiaIntervals = Count(se.size());
    e.fSandiaEnergies = Copy(se);
    e.fSandiaCoefficients = Copy(sc);
    slot = e;
  });
}

// :466-566 (e-/e+ data)
void ReadElectron(In& in, G4HepEmElectronData*& out) {
  if (in.null()) return;
  out = Zeroed<G4HepEmElectronData>(1);
  G4HepEmElectronData* d = out;
  std::vector<double> grid, loss, xs, nucGrid, nuc, tr1, ioni, sb, rb;
  std::vector<int> xsStart, ioniStart, sbStart, rbStart;
  Keys ks{"fNumMatCuts", "fNumMaterials", "fELossLogMinEkin", "fELossEILDelta", "fELossEnergyGrid", "fELossData",
          "fResMacXSecStartIndexPerMatCut", "fResMacXSecData", "fENucLogMinEkin", "fENucEILDelta", "fENucEnergyGrid",
          "fENucMacXsecData", "fTr1MacXSecData", "fElemSelectorIoniStartIndexPerMatCut", "fElemSelectorIoniData",
          "fElemSelectorBremSBStartIndexPerMatCut", "fElemSelectorBremSBData", "fElemSelectorBremRBStartIndexPerMatCut",
          "fElemSelectorBremRBData"};
  in.object([&](const std::string& k) {
    if (!ks.mark(k)) return false;
    if (k == "fNumMatCuts") d->fNumMatCuts = in.inum();
    else if (k == "fNumMaterials") d->fNumMaterials = in.inum();
    else if (k == "fELossLogMinEkin") d->fELossLogMinEkin = in.num();
    else if (k == "fELossEILDelta") d->fELossEILDelta = in.num();
    else if (k == "fELossEnergyGrid") grid = in.arr<double>();
    else if (k == "fELossData") loss = in.arr<double>();
    else if (k == "fResMacXSecStartIndexPerMatCut") xsStart = in.arr<int>();
    else if (k == "fResMacXSecData") xs = in.arr<double>();
    else if (k == "fENucLogMinEkin") d->fENucLogMinEkin = in.num();
    else if (k == "fENucEILDelta") d->fENucEILDelta = in.num();
    else if (k == "fENucEnergyGrid") nucGrid = in.arr<double>();
    else if (k == "fENucMacXsecData") nuc = in.arr<double>();
    else if (k == "fTr1MacXSecData") tr1 = in.arr<double>();
    else if (k == "fElemSelectorIoniStartIndexPerMatCut") ioniStart = in.arr<int>();
    else if (k == "fElemSelectorIoniData") ioni = in.arr<double>();
    else if (k == "fElemSelectorBremSBStartIndexPerMatCut") sbStart = in.arr<int>();
    else if (k == "fElemSelectorBremSBData") sb = in.arr<double>();
    else if (k == "fElemSelectorBremRBStartIndexPerMatCut") rbStart = in.arr<int>();
    else rb = in.arr<double>();
    return true;
  });
  ks.require();
  const size_t nmc = (size_t)d->fNumMatCuts, nmat = (size_t)d->fNumMaterials;
  const size_t N = grid.size(), NN = nucGrid.size();
  // the sizes the reference's reader notes as to-be-validated (:500-560)
  Check(d->fNumMatCuts > 0 && d->fNumMaterials > 0 && N >= 2 && NN >= 2);
  Check(loss.size() == 5 * N * nmc && xsStart.size() == nmc && nuc.size() == 2 * NN * nmat && tr1.size() == 2 * N * nmat);
  Check(ioniStart.size() == nmc && sbStart.size() == nmc && rbStart.size() == nmc);
  for (int s : xsStart) Check(s >= 0 && (size_t)s < xs.size());
  d->fELossEnergyGridSize = Count(N);
  d->fELossEnergyGrid = Copy(grid);
  d->fELossData = Copy(loss);
  d->fResMacXSecNumData = Count(xs.size());
  d->fResMacXSecStartIndexPerMatCut = Copy(xsStart);
  d->fResMacXSecData = Copy(xs);
  d->fENucEnergyGridSize = Count(NN);
  d->fENucEnergyGrid = Copy(nucGrid);
  d->fENucMacXsecData = Copy(nuc);
  d->fTr1MacXSecData = Copy(tr1);
  d->fElemSelectorIoniNumData = Count(ioni.size());
  d->fElemSelectorIoniStartIndexPerMatCut = Copy(ioniStart);
  d->fElemSelectorIoniData = Copy(ioni);
  d->fElemSelectorBremSBNumData = Count(sb.size());
  d->fElemSelectorBremSBStartIndexPerMatCut = Copy(sbStart);
  d->fElemSelectorBremSBData = Copy(sb);
  d->fElemSelectorBremRBNumData = Count(rb.size());
  d->fElemSelectorBremRBStartIndexPerMatCut = Copy(rbStart);
  d->fElemSelectorBremRBData = Copy(rb);
}

// :702-773 (Seltzer-Berger tables; the dimensions 99/65/54 are fixed)
void ReadSB(In& in, G4HepEmSBTableData*& out) {
  if (in.null()) return;
  out = Zeroed<G4HepEmSBTableData>(1);
  G4HepEmSBTableData* d = out;
  d->fMaxZet = 99;
  d->fNumElEnergy = 65;
  d->fNumKappa = 54;
  std::vector<int> start, cut;
  std::vector<double> data;
  Keys ks{"fLogMinElEnergy", "fILDeltaElEnergy", "fElEnergyVect", "fLElEnergyVect", "fKappaVect", "fLKappaVect",
          "fGammaCutIndxStartIndexPerMC", "fGammaCutIndices", "fSBStartTablesStartPerZ", "fSBTableData"};
  in.object([&](const std::string& k) {
    if (!ks.mark(k)) return false;
    if (k == "fLogMinElEnergy") d->fLogMinElEnergy = in.num();
    else if (k == "fILDeltaElEnergy") d->fILDeltaElEnergy = in.num();
    else if (k == "fElEnergyVect") in.fixed(d->fElEnergyVect);
    else if (k == "fLElEnergyVect") in.fixed(d->fLElEnergyVect);
    else if (k == "fKappaVect") in.fixed(d->fKappaVect);
    else if (k == "fLKappaVect") in.fixed(d->fLKappaVect);
    else if (k == "fGammaCutIndxStartIndexPerMC") start = in.arr<int>();
    else if (k == "fGammaCutIndices") cut = in.arr<int>();
    else if (k == "fSBStartTablesStartPerZ") in.fixed(d->fSBTablesStartPerZ);
    else data = in.arr<double>();
    return true;
  });
  ks.require();
  for (int s : d->fSBTablesStartPerZ) Check(s < 0 || (size_t)s < data.size());
  d->fNumHepEmMatCuts = Count(start.size());
  d->fNumElemsInMatCuts = Count(cut.size());
  d->fNumSBTableData = Count(data.size());
  d->fGammaCutIndxStartIndexPerMC = Copy(start);
  d->fGammaCutIndices = Copy(cut);
  d->fSBTableData = Copy(data);
}

// :780-894 (gamma data; the grid sizes 32/32/256 are fixed)
void ReadGamma(In& in, G4HepEmGammaData*& out) {
  if (in.null()) return;
  out = Zeroed<G4HepEmGammaData>(1);
  G4HepEmGammaData* d = out;
  d->fEGridSize0 = 32;
  d->fEGridSize1 = 32;
  d->fEGridSize2 = 256;
  std::vector<double> mx, cgrid, cdata;
  std::vector<int> cstart;
  Keys ks{"fNumMaterials", "fDataPerMat", "fNumData0", "fNumData1", "fEMin0", "fEMax0", "fLogEMin0", "fEILDelta0",
          "fEMax1", "fLogEMin1", "fEILDelta1", "fEMax2", "fLogEMin2", "fEILDelta2", "fMacXsecData",
          "fElemSelectorConvLogMinEkin", "fElemSelectorConvEILDelta", "fElemSelectorConvStartIndexPerMat",
          "fElemSelectorConvEgrid", "fElemSelectorConvData"};
  in.object([&](const std::string& k) {
    if (!ks.mark(k)) return false;
    if (k == "fNumMaterials") d->fNumMaterials = in.inum();
    else if (k == "fDataPerMat") d->fDataPerMat = in.inum();
    else if (k == "fNumData0") d->fNumData0 = in.inum();
    else if (k == "fNumData1") d->fNumData1 = in.inum();
    else if (k == "fEMin0") d->fEMin0 = in.num();
    else if (k == "fEMax0") d->fEMax0 = in.num();
    else if (k == "fLogEMin0") d->fLogEMin0 = in.num();
    else if (k == "fEILDelta0") d->fEILDelta0 = in.num();
    else if (k == "fEMax1") d->fEMax1 = in.num();
    else if (k == "fLogEMin1") d->fLogEMin1 = in.num();
    else if (k == "fEILDelta1") d->fEILDelta1 = in.num();
    else if (k == "fEMax2") d->fEMax2 = in.num();
    else if (k == "fLogEMin2") d->fLogEMin2 = in.num();
    else if (k == "fEILDelta2") d->fEILDelta2 = in.num();
    else if (k == "fMacXsecData") mx = in.arr<double>();
    else if (k == "fElemSelectorConvLogMinEkin") d->fElemSelectorConvLogMinEkin = in.num();
    else if (k == "fElemSelectorConvEILDelta") d->fElemSelectorConvEILDelta = in.num();
    else if (k == "fElemSelectorConvStartIndexPerMat") cstart = in.arr<int>();
    else if (k == "fElemSelectorConvEgrid") cgrid = in.arr<double>();
    else cdata = in.arr<double>();
    return true;
  });
  ks.require();
  Check(d->fNumMaterials > 0 && d->fDataPerMat > 0 && mx.size() == (size_t)d->fNumMaterials * (size_t)d->fDataPerMat);
  Check(cstart.size() == (size_t)d->fNumMaterials);
  d->fMacXsecData = Copy(mx);
  d->fElemSelectorConvStartIndexPerMat = Copy(cstart);
  d->fElemSelectorConvEgridSize = Count(cgrid.size());
  d->fElemSelectorConvEgrid = Copy(cgrid);
  d->fElemSelectorConvNumData = Count(cdata.size());
  d->fElemSelectorConvData = Copy(cdata);
}

// :899-935 (G4HepEmData)
G4HepEmData* ReadData(In& in) {
  if (in.null()) return nullptr;
  G4HepEmData* d = Zeroed<G4HepEmData>(1);
  try {
    Keys ks{"fTheMatCutData", "fTheMaterialData", "fTheElementData", "fTheElectronData", "fThePositronData",
            "fTheSBTableData", "fTheGammaData"};
    in.object([&](const std::string& k) {
      if (!ks.mark(k)) return false;
      if (k == "fTheMatCutData") ReadMatCut(in, d->fTheMatCutData);
      else if (k == "fTheMaterialData") ReadMaterials(in, d->fTheMaterialData);
      else if (k == "fTheElementData") ReadElements(in, d->fTheElementData);
      else if (k == "fTheElectronData") ReadElectron(in, d->fTheElectronData);
      else if (k == "fThePositronData") ReadElectron(in, d->fThePositronData);
      else if (k == "fTheSBTableData") ReadSB(in, d->fTheSBTableData);
      else ReadGamma(in, d->fTheGammaData);
      return true;
    });
    ks.require();
    Check(d->fTheMatCutData && d->fTheMaterialData && d->fTheElementData && d->fTheElectronData && d->fThePositronData &&
          d->fTheSBTableData && d->fTheGammaData);
    Check(d->fTheElectronData->fNumMatCuts == d->fTheMatCutData->fNumMatCutData &&
          d->fThePositronData->fNumMatCuts == d->fTheMatCutData->fNumMatCutData &&
          d->fTheGammaData->fNumMaterials == d->fTheMaterialData->fNumMaterialData);
  } catch (...) {
    hepem_tables_free(d, nullptr);
    throw;
  }
  return d;
}

// ---- writing -------------------------------------------------------------------------------------

class Out {
 public:
  explicit Out(char* buf) : buf_(buf) {}
  uint64_t size() const { return n_; }
  bool ok() const { return ok_; }

  void raw(const char* s, size_t n) {
    if (buf_) std::memcpy(buf_ + n_, s, n);
    n_ += n;
  }
  void raw(const char* s) { raw(s, std::strlen(s)); }
  void num(double v) {
    if (!std::isfinite(v)) {  // JSON has no NaN/Inf (the reference's writer would emit null, unreadable)
      ok_ = false;
      raw("null");
      return;
    }
    char b[40];
    const int k = std::snprintf(b, sizeof b, "%.17g", v);
    raw(b, (size_t)k);
  }
  void num(int v) {
    char b[16];
    const int k = std::snprintf(b, sizeof b, "%d", v);
    raw(b, (size_t)k);
  }
  void boolean(bool v) { raw(v ? "true" : "false"); }

  // members of an object, separated by commas
  void key(const char* k) {
    if (!first_) raw(",");
    first_ = false;
    raw("\"");
    raw(k);
    raw("\":");
  }
  void open() {
    raw("{");
    first_ = true;
  }
  void close() {
    raw("}");
    first_ = false;
  }
  template <typename T> void arr(const T* v, int64_t n) {
    if (n <= 0 || !v) {  // the reference writes empty arrays as null (:84-99)
      raw("null");
      return;
    }
    raw("[");
    for (int64_t i = 0; i < n; ++i) {
      if (i) raw(",");
      num(v[i]);
    }
    raw("]");
  }
  template <typename T> void kv(const char* k, T v) {
    key(k);
    num(v);
  }
  void kb(const char* k, bool v) {
    key(k);
    boolean(v);
  }
  template <typename T> void ka(const char* k, const T* v, int64_t n) {
    key(k);
    arr(v, n);
  }
  // an array whose elements are written by f(i)
  template <typename F> void objects(const char* k, int64_t n, F f) {
    key(k);
    if (n <= 0) {
      raw("null");
      return;
    }
    raw("[");
    for (int64_t i = 0; i < n; ++i) {
      if (i) raw(",");
      f(i);
    }
    raw("]");
    first_ = false;
  }

 private:
  char* buf_;
  uint64_t n_ = 0;
  bool first_ = true;
  bool ok_ = true;
};

void WriteParameters(Out& o, const G4HepEmParameters* p) {
  o.open();
  o.kv("fElectronTrackingCut", p->fElectronTrackingCut);
  o.kv("fGammaTrackingCut", p->fGammaTrackingCut);
  o.kv("fMinLossTableEnergy", p->fMinLossTableEnergy);
  o.kv("fMaxLossTableEnergy", p->fMaxLossTableEnergy);
  o.kv("fNumLossTableBins", p->fNumLossTableBins);
  o.kv("fElectronBremModelLim", p->fElectronBremModelLim);
  o.kb("fIsMSCPositronCor", p->fIsMSCPositronCor);
  o.kb("fIsMSCDisplacement", p->fIsMSCDisplacement);
  o.kv("fNumRegions", p->fNumRegions);
  o.objects("fParametersPerRegion", p->fNumRegions, [&](int64_t i) {
    const G4HepEmRegionParmeters& r = p->fParametersPerRegion[i];
    o.open();
    o.kv("fFinalRange", r.fFinalRange);
    o.kv("fDRoverRange", r.fDRoverRange);
    o.kv("fLinELossLimit", r.fLinELossLimit);
    o.kv("fMSCRangeFactor", r.fMSCRangeFactor);
    o.kv("fMSCSafetyFactor", r.fMSCSafetyFactor);
    o.kb("fIsMSCMinimalStepLimit", r.fIsMSCMinimalStepLimit);
    o.kb("fIsELossFluctuation", r.fIsELossFluctuation);
    o.kb("fIsMultipleStepsInMSCTrans", r.fIsMultipleStepsInMSCTrans);
    o.kb("fIsApplyCuts", r.fIsApplyCuts);
    o.close();
  });
  o.close();
}

void WriteElectron(Out& o, const G4HepEmElectronData* d) {
  const int64_t N = d->fELossEnergyGridSize, NN = d->fENucEnergyGridSize;
  o.open();
  o.kv("fNumMatCuts", d->fNumMatCuts);
  o.kv("fNumMaterials", d->fNumMaterials);
  o.kv("fELossLogMinEkin", d->fELossLogMinEkin);
  o.kv("fELossEILDelta", d->fELossEILDelta);
  o.ka("fELossEnergyGrid", d->fELossEnergyGrid, N);
  o.ka("fELossData", d->fELossData, 5 * N * d->fNumMatCuts);
  o.ka("fResMacXSecStartIndexPerMatCut", d->fResMacXSecStartIndexPerMatCut, d->fNumMatCuts);
  o.ka("fResMacXSecData", d->fResMacXSecData, d->fResMacXSecNumData);
  o.kv("fENucLogMinEkin", d->fENucLogMinEkin);
  o.kv("fENucEILDelta", d->fENucEILDelta);
  o.ka("fENucEnergyGrid", d->fENucEnergyGrid, NN);
  o.ka("fENucMacXsecData", d->fENucMacXsecData, 2 * NN * d->fNumMaterials);
  o.ka("fTr1MacXSecData", d->fTr1MacXSecData, 2 * N * d->fNumMaterials);
  o.ka("fElemSelectorIoniStartIndexPerMatCut", d->fElemSelectorIoniStartIndexPerMatCut, d->fNumMatCuts);
  o.ka("fElemSelectorIoniData", d->fElemSelectorIoniData, d->fElemSelectorIoniNumData);
  o.ka("fElemSelectorBremSBStartIndexPerMatCut", d->fElemSelectorBremSBStartIndexPerMatCut, d->fNumMatCuts);
  o.ka("fElemSelectorBremSBData", d->fElemSelectorBremSBData, d->fElemSelectorBremSBNumData);
  o.ka("fElemSelectorBremRBStartIndexPerMatCut", d->fElemSelectorBremRBStartIndexPerMatCut, d->fNumMatCuts);
  o.ka("fElemSelectorBremRBData", d->fElemSelectorBremRBData, d->fElemSelectorBremRBNumData);
  o.close();
}

void WriteData(Out& o, const G4HepEmData* d) {
  o.open();
  const G4HepEmMatCutData* mc = d->fTheMatCutData;
  o.key("fTheMatCutData");
  o.open();
  o.kv("fNumG4MatCuts", mc->fNumG4MatCuts);
  o.kv("fNumMatCutData", mc->fNumMatCutData);
  o.ka("fG4MCIndexToHepEmMCIndex", mc->fG4MCIndexToHepEmMCIndex, mc->fNumG4MatCuts);
  o.objects("fMatCutData", mc->fNumMatCutData, [&](int64_t i) {
    const G4HepEmMCCData& c = mc->fMatCutData[i];
    o.open();
    o.kv("fSecElProdCutE", c.fSecElProdCutE);
    o.kv("fSecPosProdCutE", c.fSecPosProdCutE);
    o.kv("fSecGamProdCutE", c.fSecGamProdCutE);
    o.kv("fLogSecGamCutE", c.fLogSecGamCutE);
    o.kv("fHepEmMatIndex", c.fHepEmMatIndex);
    o.kv("fG4MatCutIndex", c.fG4MatCutIndex);
    o.kv("fG4RegionIndex", c.fG4RegionIndex);
    o.close();
  });
  o.close();

  const G4HepEmMaterialData* md = d->fTheMaterialData;
  o.key("fTheMaterialData");
  o.open();
  o.kv("fNumG4Material", md->fNumG4Material);
  o.kv("fNumMaterialData", md->fNumMaterialData);
  o.ka("fG4MatIndexToHepEmMatIndex", md->fG4MatIndexToHepEmMatIndex, md->fNumG4Material);
  o.objects("fMaterialData", md->fNumMaterialData, [&](int64_t i) {
    const G4HepEmMatData& m = md->fMaterialData[i];
    o.open();
    o.kv("fG4MatIndex", m.fG4MatIndex);
    o.ka("fElementVect", m.fElementVect, m.fNumOfElement);
    o.ka("fNumOfAtomsPerVolumeVect", m.fNumOfAtomsPerVolumeVect, m.fNumOfElement);
    o.kv("fDensity", m.fDensity);
    o.kv("fDensityCorfactor", m.fDensityCorFactor);
    o.kv("fElectronDensity", m.fElectronDensity);
    o.kv("fRadiationLength", m.fRadiationLength);
    o.kv("fMeanExEnergy", m.fMeanExEnergy);
    o.ka("fSandiaEnergies", m.fSandiaEnergies, m.fNumOfSandiaIntervals);
    o.ka("fSandiaCoefficients", m.fSandiaCoefficients, 4 * (int64_t)m.fNumOfSandiaIntervals);
    o.kv("fZeff", m.fZeff);
    o.kv("fZeff23", m.fZeff23);
    o.kv("fZeffSqrt", m.fZeffSqrt);
    o.kv("fUMSCPar", m.fUMSCPar);
    o.ka("fUMSCStepMinPars", m.fUMSCStepMinPars, 2);
    o.ka("fUMSCTailCoeff", m.fUMSCTailCoeff, 4);
    o.ka("fUMSCThetaCoeff", m.fUMSCThetaCoeff, 2);
    o.close();
  });
  o.close();

  // only the elements in use (fZet > 0), as the reference persists them
  const G4HepEmElementData* ed = d->fTheElementData;
  std::vector<int> zs;
  for (int z = 0; z <= ed->fMaxZet; ++z)
    if (ed->fElementData[z].fZet > 0.0) zs.push_back(z);
  o.objects("fTheElementData", (int64_t)zs.size(), [&](int64_t i) {
    const G4HepEmElemData& e = ed->fElementData[zs[(size_t)i]];
    o.open();
    o.kv("fZet", e.fZet);
    o.kv("fZet13", e.fZet13);
    o.kv("fZet23", e.fZet23);
    o.kv("fCoulomb", e.fCoulomb);
    o.kv("fLogZ", e.fLogZ);
    o.kv("fZFactor1", e.fZFactor1);
    o.kv("fDeltaMaxLow", e.fDeltaMaxLow);
    o.kv("fDeltaMaxHigh", e.fDeltaMaxHigh);
    o.kv("fILVarS1", e.fILVarS1);
    o.kv("fILVarS1Cond", e.fILVarS1Cond);
    o.ka("fSandiaEnergies", e.fSandiaEnergies, e.fNumOfSandiaIntervals);
    o.ka("fSandiaCoefficients", e.fSandiaCoefficients, 4 * (int64_t)e.fNumOfSandiaIntervals);
    o.kv("fKShellBindingEnergy", e.fKShellBindingEnergy);
    o.close();
  });

  o.key("fTheElectronData");
  WriteElectron(o, d->fTheElectronData);
  o.key("fThePositronData");
  WriteElectron(o, d->fThePositronData);

  const G4HepEmSBTableData* sb = d->fTheSBTableData;
  o.key("fTheSBTableData");
  o.open();
  o.kv("fLogMinElEnergy", sb->fLogMinElEnergy);
  o.kv("fILDeltaElEnergy", sb->fILDeltaElEnergy);
  o.ka("fElEnergyVect", sb->fElEnergyVect, 65);
  o.ka("fLElEnergyVect", sb->fLElEnergyVect, 65);
  o.ka("fKappaVect", sb->fKappaVect, 54);
  o.ka("fLKappaVect", sb->fLKappaVect, 54);
  o.ka("fGammaCutIndxStartIndexPerMC", sb->fGammaCutIndxStartIndexPerMC, sb->fNumHepEmMatCuts);
  o.ka("fGammaCutIndices", sb->fGammaCutIndices, sb->fNumElemsInMatCuts);
  o.ka("fSBStartTablesStartPerZ", sb->fSBTablesStartPerZ, 121);
  o.ka("fSBTableData", sb->fSBTableData, sb->fNumSBTableData);
  o.close();

  const G4HepEmGammaData* g = d->fTheGammaData;
  o.key("fTheGammaData");
  o.open();
  o.kv("fNumMaterials", g->fNumMaterials);
  o.kv("fDataPerMat", g->fDataPerMat);
  o.kv("fNumData0", g->fNumData0);
  o.kv("fNumData1", g->fNumData1);
  o.kv("fEMin0", g->fEMin0);
  o.kv("fEMax0", g->fEMax0);
  o.kv("fLogEMin0", g->fLogEMin0);
  o.kv("fEILDelta0", g->fEILDelta0);
  o.kv("fEMax1", g->fEMax1);
  o.kv("fLogEMin1", g->fLogEMin1);
  o.kv("fEILDelta1", g->fEILDelta1);
  o.kv("fEMax2", g->fEMax2);
  o.kv("fLogEMin2", g->fLogEMin2);
  o.kv("fEILDelta2", g->fEILDelta2);
  o.ka("fMacXsecData", g->fMacXsecData, (int64_t)g->fNumMaterials * g->fDataPerMat);
  o.kv("fElemSelectorConvLogMinEkin", g->fElemSelectorConvLogMinEkin);
  o.kv("fElemSelectorConvEILDelta", g->fElemSelectorConvEILDelta);
  o.ka("fElemSelectorConvStartIndexPerMat", g->fElemSelectorConvStartIndexPerMat, g->fNumMaterials);
  o.ka("fElemSelectorConvEgrid", g->fElemSelectorConvEgrid, g->fElemSelectorConvEgridSize);
  o.ka("fElemSelectorConvData", g->fElemSelectorConvData, g->fElemSelectorConvNumData);
  o.close();
  o.close();
}

bool DataComplete(const G4HepEmData* d) {
  return d && d->fTheMatCutData && d->fTheMaterialData && d->fTheElementData && d->fTheElectronData &&
         d->fThePositronData && d->fTheSBTableData && d->fTheGammaData;
}

}  // namespace

extern "C" int hepem_tables_from_json(const char* text, uint64_t len, G4HepEmData** out_data, G4HepEmParameters** out_pars) {
  if (!text || !out_data || !out_pars) return HEPEM_ERR_INVALID;
  *out_data = nullptr;
  *out_pars = nullptr;
  G4HepEmData* data = nullptr;
  G4HepEmParameters* pars = nullptr;
  try {
    In in(text, (size_t)len);
    // a G4HepEmState document, or a G4HepEmData one, or a G4HepEmParameters one
    bool state = false, any = false;
    in.object([&](const std::string& k) {
      if (k == "fParameters" || k == "fData") {
        state = true;
        if (k == "fParameters") {
          if (pars) throw BadJson{};
          pars = ReadParameters(in);
        } else {
          if (data) throw BadJson{};
          data = ReadData(in);
        }
        any = true;
        return true;
      }
      return false;
    });
    if (!in.atEnd() || !state || !any || !data || !pars) throw BadJson{};
    // every index the tables hold must point inside them
    if (!HepemTablesValid(data)) throw BadJson{};
  } catch (...) {
    hepem_tables_free(data, pars);
    return HEPEM_ERR_INVALID;
  }
  *out_data = data;
  *out_pars = pars;
  return HEPEM_OK;
}

namespace {
void WriteState(Out& o, const G4HepEmData* data, const G4HepEmParameters* pars) {
  o.open();
  o.key("fData");
  WriteData(o, data);
  o.key("fParameters");
  WriteParameters(o, pars);
  o.close();
}
}  // namespace

extern "C" int hepem_tables_to_json(const G4HepEmData* data, const G4HepEmParameters* pars, char* buf, uint64_t* size) {
  if (!size || !DataComplete(data) || !pars || pars->fNumRegions < 1 || !pars->fParametersPerRegion)
    return HEPEM_ERR_INVALID;
  Out sizing(nullptr);  // writes no byte: measures the text
  WriteState(sizing, data, pars);
  if (!sizing.ok()) return HEPEM_ERR_INVALID;
  if (!buf) {
    *size = sizing.size();
    return HEPEM_OK;
  }
  if (*size < sizing.size()) return HEPEM_ERR_CAPACITY;
  Out o(buf);
  WriteState(o, data, pars);
  *size = o.size();
  return HEPEM_OK;
}
