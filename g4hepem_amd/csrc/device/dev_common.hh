// dev_common.hh -- device math, the per-track Philox stream and the flattened
// device table image used by every kernel.
//
// All device code is compiled with -ffp-contract=off: in the FP64 form every
// product and sum is rounded as written, which is what makes the FP64 results
// reproduce the host reference's arithmetic bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __device__ __forceinline__

namespace hepem {

// ---- constants (the reference's G4HepEmConstants.hh:7-29 values) ---------------------
constexpr double kPi = 3.1415926535897931e+00;
constexpr double k2Pi = 2.0 * 3.1415926535897931e+00;
constexpr double kElectronMassC2 = 5.1099890999999997e-01;
constexpr double kInvElectronMassC2 = 1.0 / 5.1099890999999997e-01;
constexpr double kAlpha = 7.2973525653052150e-03;
constexpr double kPir02 = 2.4946724123674787e-23;
constexpr double kMigdalConst = 5.2804955733859579e-30;
constexpr double kLPMconstant = 7.6843819381368661e+05;
constexpr double kALargeValue = 1.0E+20;
constexpr int kPoissonMaxTerms = 1000;
constexpr int kSBGuide = 32;           // buckets of the SB search guide

template <typename T> HD T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> HD T tmin(T a, T b) { return a < b ? a : b; }

// ---- exp / log --------------------------------------------------------------------------
// FP64: the Cephes rational forms (S. L. Moshier's Cephes library, as packaged
// in VDT and used on the reference host path, G4HepEmExp.hh / G4HepEmLog.hh),
// evaluated in the same order so the device reproduces the host values.
HD double vexp(double initial_x) {
  double x = initial_x;
  const double t = 1.4426950408889634073599 * x + 0.5;
  int32_t fl = (int32_t)t;
  fl -= (int32_t)(__float_as_uint((float)t) >> 31);
  double px = (double)fl;
  const int32_t n = (int32_t)px;
  x -= px * 6.93145751953125E-1;
  x -= px * 1.42860682030941723212E-6;
  const double xx = x * x;
  px = 1.26177193074810590878E-4;
  px *= xx;
  px += 3.02994407707441961300E-2;
  px *= xx;
  px += 9.99999999999999999910E-1;
  px *= x;
  double qx = 3.00198505138664455042E-6;
  qx *= xx;
  qx += 2.52448340349684104192E-3;
  qx *= xx;
  qx += 2.27265548208155028766E-1;
  qx *= xx;
  qx += 2.00000000000000000009E0;
  x = px / (qx - px);
  x = 1.0 + 2.0 * x;
  x *= __longlong_as_double((long long)((((uint64_t)(int64_t)n) + 1023) << 52));
  if (initial_x > 708.0) x = __longlong_as_double(0x7FF0000000000000LL);
  if (initial_x < -708.0) x = 0.;
  return x;
}

HD double vlog(double x) {
  const double original_x = x;
  uint64_t n = (uint64_t)__double_as_longlong(x);
  const int32_t e = (int32_t)(n >> 52);
  double fe = e - 1023;
  n &= 0x800FFFFFFFFFFFFFULL;
  n |= 0x3FE0000000000000ULL;
  x = __longlong_as_double((long long)n);
  if (x > 0.70710678118654752440) fe += 1.; else x += x;
  x -= 1.0;
  double px = 1.01875663804580931796E-4;
  px *= x; px += 4.97494994976747001425E-1;
  px *= x; px += 4.70579119878881725854E0;
  px *= x; px += 1.44989225341610930846E1;
  px *= x; px += 1.79368678507819816313E1;
  px *= x; px += 7.70838733755885391666E0;
  const double x2 = x * x;
  px *= x;
  px *= x2;
  double qx = x;
  qx += 1.12873587189167450590E1;
  qx *= x; qx += 4.52279145837532221105E1;
  qx *= x; qx += 8.29875266912776603211E1;
  qx *= x; qx += 7.11544750618563894466E1;
  qx *= x; qx += 2.31251620126765340583E1;
  double res = px / qx;
  res -= fe * 2.121944400546905827679e-4;
  res -= 0.5 * x2;
  res = x + res;
  res += fe * 0.693359375;
  if (original_x > 1e307) res = __longlong_as_double(0x7FF0000000000000LL);
  if (original_x < 0) res = __longlong_as_double(0xFFF8000000000000LL);
  return res;
}

// FP32: hardware-accelerated single-precision transcendental functions.
HD float vexp(float x) { return __expf(x); }
HD float vlog(float x) { return __logf(x); }

template <typename R> HD R vpow(R x, R a) { return vexp(a * vlog(x)); }

HD double dsqrt(double x) { return __builtin_sqrt(x); }
HD float dsqrt(float x) { return __builtin_sqrtf(x); }
HD double dsin(double x) { return sin(x); }
HD float dsin(float x) { return __sinf(x); }
HD double dcos(double x) { return cos(x); }
HD float dcos(float x) { return __cosf(x); }
// sine and cosine of one angle: one argument reduction on the device
HD void dsincos(double x, double& s, double& c) {
#if defined(__HIP_DEVICE_COMPILE__)
  sincos(x, &s, &c);
#else
  s = sin(x);
  c = cos(x);
#endif
}
HD void dsincos(float x, float& s, float& c) {
  s = __sinf(x);
  c = __cosf(x);
}
HD double dabs(double x) { return __builtin_fabs(x); }
HD float dabs(float x) { return __builtin_fabsf(x); }

// ---- per-track random stream ------------------------------------------------------------
// Philox4x32-10 (J. Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC'11). Draw d of track `id` comes from block counter {d/2, id, 0, 0}; double j
// of a block takes words (2j, 2j+1) and maps their top 52 bits k to (k+0.5)*2^-52.
HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[1] = lo1;
    c[3] = lo0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct Rng {
  uint32_t k0, k1, id, hi, ctr;
  uint32_t blk;          // block held in buf (0xFFFFFFFF: none)
  double buf[2];
  bool has_gauss;        // G4HepEmRandomEngine::fIsGauss
  double gauss;          // G4HepEmRandomEngine::fGauss

  HD void init(uint64_t seed, uint32_t tid, uint32_t thi, uint32_t c, bool hg, double g) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32);
    id = tid;
    hi = thi;
    ctr = c;
    blk = 0xFFFFFFFFu;
    has_gauss = hg;
    gauss = g;
  }
  HD double flat() {
    const uint32_t d = ctr++;
    const uint32_t b = d >> 1;
    if (b != blk) {
      uint32_t c[4] = {b, id, hi, 0u};
      philox4x32_10(c, k0, k1);
      const uint64_t w0 = ((uint64_t)c[1] << 32) | c[0];
      const uint64_t w1 = ((uint64_t)c[3] << 32) | c[2];
      buf[0] = ((double)(w0 >> 12) + 0.5) * 2.220446049250313080847e-16;
      buf[1] = ((double)(w1 >> 12) + 0.5) * 2.220446049250313080847e-16;
      blk = b;
    }
    return (d & 1u) ? buf[1] : buf[0];
  }
  template <typename R> HD void flatArray(int n, R* v) {
    for (int i = 0; i < n; ++i) v[i] = (R)flat();
  }
  // G4HepEmRandomEngine::Gauss (G4HepEmRandomEngine.hh:50-67): polar Marsaglia
  // with the second deviate cached in the track state.
  template <typename R> HD R Gauss(R mean, R stDev) {
    if (has_gauss) {
      has_gauss = false;
      return (R)gauss * stDev + mean;
    }
    R rnd[2], r, v1, v2;
    do {
      flatArray(2, rnd);
      v1 = (R)2. * rnd[0] - (R)1.;
      v2 = (R)2. * rnd[1] - (R)1.;
      r = v1 * v1 + v2 * v2;
    } while (r > (R)1.);
    const R fac = dsqrt((R)-2. * vlog(r) / r);
    gauss = (double)(v1 * fac);
    has_gauss = true;
    return v2 * fac * stDev + mean;
  }
  // G4HepEmRandomEngine::Poisson (G4HepEmRandomEngine.hh:73-97)
  template <typename R> HD int Poisson(R mean) {
    const int border = 16;
    const R limit = (R)2.E+9;
    int number = 0;
    if (mean <= (R)border) {
      const R position = (R)flat();
      R poissonValue = vexp(-mean);
      R poissonSum = poissonValue;
      // kPoissonMaxTerms bounds the summation; for mean <= 16 the sum reaches 1
      // within double rounding long before it, so the bound can only act where
      // the unbounded loop would never end (a sum saturating below `position`)
      while (poissonSum <= position && number < kPoissonMaxTerms) {
        ++number;
        poissonValue *= mean / (R)number;
        poissonSum += poissonValue;
      }
      return number;
    }
    R rnd[2];
    flatArray(2, rnd);
    const R t = dsqrt((R)-2. * vlog(rnd[0])) * dcos((R)k2Pi * rnd[1]);
    const R value = mean + t * dsqrt(mean) + (R)0.5;
    return value < (R)0. ? 0 : value >= limit ? (int)limit : (int)value;
  }
};

// ---- flattened table image ---------------------------------------------------------------
// One device allocation holds every array; this struct, passed by value as a
// kernel argument, carries the array bases (in SGPRs) and the scalars. Arrays
// of table values use the precision R of the form (double or float); per-couple,
// per-material and per-element scalars stay double.
// Abscissas (energy, range and selector grids) are read from the *64 arrays,
// which are double in both forms: the interval fraction (x - x1)/(x2 - x1) of an
// interpolation loses all precision to cancellation when x and x1 are rounded
// to float near a grid node. In the FP64 form they alias the R arrays.
// One restricted macroscopic cross-section block (ioni or brem of one
// couple) of G4HepEmElectronData::fResMacXSecData with its header unpacked, so
// a lookup reads one descriptor and then the [E, xs, xs''] triples at
// xsData[start] instead of following the header chain.
template <typename R>
struct XsBlock {
  double minE, lastE;   // first and last grid energy
  R logMin, ild;        // log of the first energy, 1/log-spacing
  int start, n;         // first triple, number of triples
  double maxE;          // energy of the largest cross section
  R maxV;               // the largest cross section
};

template <typename R>
struct DevEl {
  const R* lossGrid;    // [N] loss energy grid
  const double* lossGrid64;
  const R* lossData;    // [5N per couple] [R,R''] [dEdx,dEdx''] [InvR'']
  const double* lossData64;
  int N;
  R lossLogMin, lossILD;
  double lossEmin, lossEmax;  // lossGrid64[0], lossGrid64[N-1]
  int rangeMonotone;    // every couple's range column is non-decreasing (enables the bin hint)
  const int* xsStart;   // [couple] start of the ioni block in xsData
  const XsBlock<R>* xsBlock;  // [2 per couple]: ioni, brem
  const R* xsData;      // restricted macroscopic cross sections
  const double* xsData64;
  int xsN;              // entries of xsData
  const R* nucGrid;     // [nucN]
  const double* nucGrid64;
  const R* nucData;     // [2 nucN per material]
  int nucN;
  R nucLogMin, nucILD;
  double nucEmin, nucEmax;    // nucGrid64[0], nucGrid64[nucN-1]
  const R* tr1;         // [2N per material]
  const int* selSBStart;
  const R* selSB;
  const double* selSB64;
  const int* selRBStart;
  const R* selRB;
  const double* selRB64;
};

template <typename R>
struct DevTables {
  DevEl<R> el[2];                 // [0] e-, [1] e+
  int nmc, nmat;
  // couples
  const double* mcElCut;
  const double* mcGamCut;
  const double* mcLogGamCut;
  const int* mcMat;
  const int* mcReg;
  // materials
  const int* matNElem;
  const int* matElemStart;       // into matElemZ / matElemNat
  const int* matElemZ;
  const double* matElemNat;
  const double* matDensCor;
  const double* matElDens;
  const double* matRadLen;
  const double* matMeanExc;
  const double* matZeff;
  const double* matZeff23;
  const double* matZeffSqrt;
  const double* matUMSCPar;
  const double* matStepMin;      // [2 per material]
  const double* matTail;         // [4 per material]
  const double* matTheta;        // [2 per material]
  const int* matSandiaStart;
  const int* matSandiaN;
  const double* matSandiaE;
  const double* matSandiaC;      // [4 per interval]
  // elements, indexed by Z (0..maxZ)
  int maxZ;
  const double* elZ13;
  const double* elZ23;
  const double* elCoulomb;
  const double* elLogZ;
  const double* elZF1;
  const double* elDMaxLow;
  const double* elDMaxHigh;
  const double* elILS1;
  const double* elILS1Cond;
  const double* elKBind;
  const int* elSandiaStart;
  const int* elSandiaN;
  const double* elSandiaE;
  const double* elSandiaC;
  // Seltzer-Berger sampling tables
  const R* sbData;
  const int* sbStartZ;           // [121]
  const int* sbGCutStart;        // [couple]
  const int* sbGCutIdx;
  const double* sbElE;           // [65]
  const double* sbLElE;          // [65]
  const R* sbLKappa;             // [54]
  double sbLogMinE, sbILDE;
  int sbNumKappa;
  // search guide of the SB sampling tables (derived at context creation):
  // row r = sbGuideStartZ[Z] + (energy index - first energy index of Z) holds
  // kSBGuide entries; entry g = largest kappa index k with cumulative value
  // cum[k] <= g / kSBGuide, in the form's precision
  const uint8_t* sbGuide;
  const int* sbGuideStartZ;      // [121], -1 for elements without a table
  // gamma macroscopic cross sections (3 windows) and the conversion selector
  const R* gmData;
  const double* gmData64;
  int gmPerMat, gmOff1, gmOff2, gmN0, gmN1, gmN2;
  double gEMax0, gEMax1;
  R gLogEMin0, gILD0, gLogEMin1, gILD1, gLogEMin2, gILD2;
  const int* convSelStart;
  const R* convSelData;
  const R* convSelGrid;
  const double* convSelGrid64;
  int convSelN;
  R convSelLogMin, convSelILD;
  // parameters
  double elTrackCut, gamTrackCut, minLossE, bremModelLim;
  int posCor, displace;
  const double* regFinalRange;
  const double* regDRover;
  const double* regLinLim;
  const double* regRangeFac;
  const double* regSafetyFac;
  const int* regMinimal;
  const int* regFluct;
  const int* regMultiStep;  // fIsMultipleStepsInMSCTrans: MSC-limited sub-steps within one step
};

}  // namespace hepem
