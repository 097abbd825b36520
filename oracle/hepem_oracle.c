/*
 * hepem_oracle.c -- TEST INFRASTRUCTURE ONLY (see hepem_oracle.h).
 *
 * Operation-by-operation C restatement of the reference's G4HepEmRun
 * run-time physics (paths relative to the reference tree, directory
 * G4HepEm/G4HepEmRun/include unless given). Compiled with
 * -ffp-contract=off so that every product and sum is rounded as written.
 *
 * The only deliberate departures from the reference are:
 *  - the random engine: G4HepEmRandomEngine::flat/flatArray are consumer
 *    supplied (G4HepEmRandomEngine.hh:22-47); here they draw consecutive
 *    doubles from a per-track Philox4x32-10 stream;
 *  - gamma SampleInteraction (G4HepEmGammaManager.icc:203-208) reads a fifth
 *    "process" beyond the stored four when the photon falls into the
 *    gamma-nuclear bin; that value is never used (the interaction is not
 *    handled), so it is not read here.
 */
#include "hepem_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- constants: G4HepEmConstants.hh:7-29 ------------------------------------ */
static const double kPi = 3.1415926535897931e+00;
static const double k2Pi = 2.0 * 3.1415926535897931e+00;
static const double kElectronMassC2 = 5.1099890999999997e-01;
static const double kInvElectronMassC2 = 1.0 / 5.1099890999999997e-01;
static const double kAlpha = 7.2973525653052150e-03;
static const double kPir02 = 2.4946724123674787e-23;
static const double kMigdalConst = 5.2804955733859579e-30;
static const double kLPMconstant = 7.6843819381368661e+05;
static const double kALargeValue = 1.0E+20;

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ---- VDT/Cephes exp and log: G4HepEmExp.hh:182-223, G4HepEmLog.hh:228-263 ----
 * The polynomial coefficients are the published Cephes ones (S. Moshier). */
static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }

double oracle_vdt_exp(double initial_x) {
  double x = initial_x;
  /* floor of LOG2E*x+0.5 as VDT computes it: truncation, minus the sign bit of
     the value rounded to float (G4HepEmExp.hh:110-114) */
  const double t = 1.4426950408889634073599 * x + 0.5;
  int32_t fl = (int32_t)t;
  fl -= (int32_t)(f2u((float)t) >> 31);
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
  x *= u2d((((uint64_t)(int64_t)n) + 1023) << 52);
  if (initial_x > 708.0) x = INFINITY;
  if (initial_x < -708.0) x = 0.;
  return x;
}

double oracle_vdt_log(double x) {
  const double original_x = x;
  uint64_t n = d2u(x);
  const int32_t e = (int32_t)(n >> 52);
  double fe = e - 1023;
  n &= 0x800FFFFFFFFFFFFFULL;
  n |= 0x3FE0000000000000ULL;
  x = u2d(n);
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
  if (original_x > 1e307) res = INFINITY;
  if (original_x < 0) res = -NAN;
  return res;
}

#define HLog(x) oracle_vdt_log(x)
#define HExp(x) oracle_vdt_exp(x)
static inline double HPow(double x, double a) { return oracle_vdt_exp(a * oracle_vdt_log(x)); }

/* ---- RNG: Philox4x32-10 stream + G4HepEmRandomEngine.hh:50-99 --------------- */
void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

typedef struct orng {
  uint32_t key[2];
  uint32_t id;
  uint32_t hi;        /* high word of the stream selector */
  uint32_t ctr;       /* draw index */
  int      has_gauss; /* G4HepEmRandomEngine::fIsGauss */
  double   gauss;     /* G4HepEmRandomEngine::fGauss   */
} orng;

/* draw d of stream (hi:id): Philox block d/2, counter {d/2, id, hi, 0}; double j
   of a block from words (2j, 2j+1): 52 high bits k -> (k + 0.5) 2^-52 in (0,1) */
static inline double rng_draw(const uint32_t key[2], uint32_t id, uint32_t hi, uint32_t d) {
  const uint32_t c[4] = {d >> 1, id, hi, 0u};
  uint32_t o[4];
  oracle_philox4x32_10(c, key, o);
  const int j = (int)(d & 1u);
  const uint64_t w = ((uint64_t)o[2 * j + 1] << 32) | (uint64_t)o[2 * j];
  return ((double)(w >> 12) + 0.5) * 2.220446049250313080847e-16;
}

void oracle_rng_uniforms(uint64_t seed, uint32_t id, uint32_t first, int n, double* out) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < n; ++i) out[i] = rng_draw(key, id, 0u, first + (uint32_t)i);
}

static inline double flat(orng* r) { return rng_draw(r->key, r->id, r->hi, r->ctr++); }
static inline void flatArray(orng* r, int n, double* v) { for (int i = 0; i < n; ++i) v[i] = flat(r); }

static double Gauss(orng* r, double mean, double stDev) {
  if (r->has_gauss) {
    r->has_gauss = 0;
    return r->gauss * stDev + mean;
  }
  double rnd[2], rr, v1, v2;
  do {
    flatArray(r, 2, rnd);
    v1 = 2. * rnd[0] - 1.;
    v2 = 2. * rnd[1] - 1.;
    rr = v1 * v1 + v2 * v2;
  } while (rr > 1.);
  const double fac = sqrt(-2. * HLog(rr) / rr);
  r->gauss = v1 * fac;
  r->has_gauss = 1;
  return v2 * fac * stDev + mean;
}

static int Poisson(orng* r, double mean) {
  const int border = 16;
  const double limit = 2.E+9;
  int number = 0;
  if (mean <= border) {
    const double position = flat(r);
    double poissonValue = HExp(-mean);
    double poissonSum = poissonValue;
    /* bounded at 1000 terms: for mean <= 16 the sum reaches 1 within double
       rounding long before, so the bound only acts where the reference's
       unbounded loop (G4HepEmRandomEngine.hh:83) would never end */
    while (poissonSum <= position && number < 1000) {
      ++number;
      poissonValue *= mean / number;
      poissonSum += poissonValue;
    }
    return number;
  }
  double rnd[2];
  flatArray(r, 2, rnd);
  const double t = sqrt(-2. * HLog(rnd[0])) * cos(k2Pi * rnd[1]);
  const double value = mean + t * sqrt(mean) + 0.5;
  return value < 0. ? 0 : value >= limit ? (int)limit : (int)value;
}

/* ---- G4HepEmRunUtils.icc ------------------------------------------------------ */
static void RotateToReferenceFrame(double* dir, const double* refDir) { /* :32-46 */
  double up = refDir[0] * refDir[0] + refDir[1] * refDir[1];
  if (up > 0.) {
    up = sqrt(up);
    const double px = dir[0], py = dir[1], pz = dir[2];
    dir[0] = (refDir[0] * refDir[2] * px - refDir[1] * py) / up + refDir[0] * pz;
    dir[1] = (refDir[1] * refDir[2] * px + refDir[0] * py) / up + refDir[1] * pz;
    dir[2] = -up * px + refDir[2] * pz;
  } else if (refDir[2] < 0.) {
    dir[0] = -dir[0];
    dir[2] = -dir[2];
  }
}

static inline double GetSpline(double x1, double x2, double y1, double y2, double sd1, double sd2, double x) { /* :49-61 */
  const double dl = x2 - x1;
  const double b = dmax(0., dmin(1., (x - x1) / dl));
  const double os = 0.166666666667;
  const double c0 = (2.0 - b) * sd1;
  const double c1 = (1.0 + b) * sd2;
  return y1 + b * (y2 - y1) + (b * (b - 1.0)) * (c0 + c1) * (dl * dl * os);
}
static inline double GetLinear(double x1, double x2, double y1, double y2, double x) { /* :63-71 */
  const double dl = x2 - x1;
  const double b = dmax(0., dmin(1., (x - x1) / dl));
  return y1 + b * (y2 - y1);
}
static inline int LogBin(double logx, double logxmin, double invLD, int ndata) {
  return (int)dmax(0., dmin((logx - logxmin) * invLD, ndata - 2.));
}
/* x grid separate, [y,y''] interleaved: :84-91 */
static inline double GetSplineLog2(int ndata, const double* xdata, const double* ydata, double x, double logx,
                                   double logxmin, double invLD) {
  const double xv = dmax(xdata[0], dmin(xdata[ndata - 1], x));
  const int idx = LogBin(logx, logxmin, invLD, ndata);
  const int idx2 = 2 * idx;
  return GetSpline(xdata[idx], xdata[idx + 1], ydata[idx2], ydata[idx2 + 2], ydata[idx2 + 1], ydata[idx2 + 3], xv);
}
/* [x,y,y''] interleaved: :95-102 */
static inline double GetSplineLog3(int ndata, const double* data, double x, double logx, double logxmin, double invLD) {
  const double xv = dmax(data[0], dmin(data[3 * (ndata - 1)], x));
  const int idx = LogBin(logx, logxmin, invLD, ndata);
  const int idx3 = 3 * idx;
  return GetSpline(data[idx3], data[idx3 + 3], data[idx3 + 1], data[idx3 + 4], data[idx3 + 2], data[idx3 + 5], xv);
}
static inline double GetLinearLog(int ndata, const double* data, double x, double logx, double logxmin, double invLD) { /* :125-133 */
  const int idx = LogBin(logx, logxmin, invLD, ndata);
  const int i0 = 2 * idx, i1 = i0 + 2;
  return GetLinear(data[i0], data[i1], data[i0 + 1], data[i1 + 1], x);
}
static inline void GetLinearLog2All(int ndata, const double* data, double x, double logx, double logxmin, double invLD,
                                    double res[2]) { /* :146-161 */
  const int idx = LogBin(logx, logxmin, invLD, ndata);
  const int i0 = 3 * idx, i1 = i0 + 3;
  const double dl = data[i1] - data[i0];
  const double b = dmax(0., dmin(1., (x - data[i0]) / dl));
  for (int i = 1; i < 3; ++i) {
    const double y1 = data[i0 + i], y2 = data[i1 + i];
    res[i - 1] = dmax(0.0, y1 + b * (y2 - y1));
  }
}
static inline double GetSplineLog4(int ndata, const double* data, double x, double logx, double logxmin, double invLD,
                                   int iwhich) { /* :178-187 */
  const int idx = LogBin(logx, logxmin, invLD, ndata);
  const int i0 = 9 * idx, i1 = i0 + 9;
  iwhich = (iwhich - 1) * 2 + 1;
  return GetSpline(data[i0], data[i1], data[i0 + iwhich], data[i1 + iwhich], data[i0 + iwhich + 1], data[i1 + iwhich + 1], x);
}
static inline int FindLowerBinIndex(const double* xdata, int num, double x, int step) { /* :227-236 */
  int ml = -1, mu = num - 1;
  while (abs(mu - ml) > 1) {
    const int mav = (int)(0.5 * (ml + mu));
    if (x < xdata[step * mav]) mu = mav; else ml = mav;
  }
  return mu - 1;
}

/* ---- track state (G4HepEmTrack.hh, G4HepEmElectronTrack.hh, G4HepEmGammaTrack.hh,
        G4HepEmMSCTrackData.hh) ------------------------------------------------ */
typedef struct otrack {
  double dir[3], ekin, lekin, charge, edep, gstep, mfp[4], nial[4], safety;
  int imc, winner, on_boundary;
} otrack;
typedef struct omsc {
  double lambtr1, true_step, zpath, displ[3], dir[3], init_range, dyn_range_factor, tlimit_min, par1, par2, par3;
  int no_scat, is_displace, first_step, active;
} omsc;
typedef struct oel { otrack t; omsc msc; double range, pstep, pre_ekin, pre_lekin; } oel;
typedef struct ogm { otrack t; double pe_mxsec; } ogm;
typedef struct osec { int ptype; double ekin, dir[3]; } osec;
typedef struct otl {
  oel el; ogm gm; orng* rng;
  int nsec_el, nsec_gm;
  osec sec_el[2], sec_gm[2];
} otl;

static inline void SetEKin(otrack* t, double e) { t->ekin = e; t->lekin = 100.0; }
static inline double GetLogEKin(otrack* t) { /* G4HepEmTrack.hh:95-100 */
  if (t->lekin > 99.0) t->lekin = (t->ekin > 0.) ? HLog(t->ekin) : -30;
  return t->lekin;
}
static inline void TrackReset(otrack* t) {
  memset(t, 0, sizeof(*t));
  t->lekin = 100.0;
  for (int i = 0; i < 4; ++i) { t->mfp[i] = -1.0; t->nial[i] = -1.0; }
  t->imc = -1; t->winner = -1; t->on_boundary = 0;
}
static inline void MscReset(omsc* m) { /* G4HepEmMSCTrackData.hh ReSet */
  memset(m, 0, sizeof(*m));
  m->dir[2] = 1.;
  m->init_range = 1.0e+21;
  m->dyn_range_factor = 0.04;
  m->tlimit_min = 1.0E-7;
  m->par1 = -1.;
  m->first_step = 1;
}
static inline osec* AddSecEl(otl* tl, int ptype) { osec* s = &tl->sec_el[tl->nsec_el++]; s->ptype = ptype; return s; }
static inline osec* AddSecGm(otl* tl) { osec* s = &tl->sec_gm[tl->nsec_gm++]; s->ptype = HEPEM_GAMMA; return s; }

/* ---- G4HepEmElectronManager.icc: table accessors ----------------------------- */
static inline const struct G4HepEmElectronData* ElData(const struct G4HepEmData* d, int isEl) {
  return isEl ? d->fTheElectronData : d->fThePositronData;
}
static double GetRestRange(const struct G4HepEmElectronData* e, int imc, double ekin, double lekin) { /* :486-493 */
  const int N = e->fELossEnergyGridSize;
  const double r = GetSplineLog2(N, e->fELossEnergyGrid, &e->fELossData[5 * N * imc], ekin, lekin, e->fELossLogMinEkin, e->fELossEILDelta);
  return dmax(0.0, r);
}
static double GetRestDEDX(const struct G4HepEmElectronData* e, int imc, double ekin, double lekin) { /* :495-502 */
  const int N = e->fELossEnergyGridSize;
  const double r = GetSplineLog2(N, e->fELossEnergyGrid, &e->fELossData[N * (5 * imc + 2)], ekin, lekin, e->fELossLogMinEkin, e->fELossEILDelta);
  return dmax(0.0, r);
}
static double GetInvRange(const struct G4HepEmElectronData* e, int imc, double range) { /* :504-520 */
  const int N = e->fELossEnergyGridSize;
  const int iR = 5 * N * imc;
  const double minRange = e->fELossData[iR];
  if (range < minRange) {
    const double dum = range / minRange;
    return dmax(0.0, e->fELossEnergyGrid[0] * dum * dum);
  }
  const int i = FindLowerBinIndex(&e->fELossData[iR], N, range, 2);
  const double* xd = &e->fELossData[iR];
  const double* sd = &e->fELossData[iR + 4 * N];
  const double en = GetSpline(xd[2 * i], xd[2 * (i + 1)], e->fELossEnergyGrid[i], e->fELossEnergyGrid[i + 1], sd[i], sd[i + 1], range);
  return dmax(0.0, en);
}
static double GetRestMacXSec(const struct G4HepEmElectronData* e, int imc, double ekin, double lekin, int isioni) { /* :522-532 */
  const int iIoni = e->fResMacXSecStartIndexPerMatCut[imc];
  const int nIoni = (int)e->fResMacXSecData[iIoni];
  const int iStart = isioni ? iIoni : iIoni + 3 * nIoni + 5;
  const int nData = (int)e->fResMacXSecData[iStart];
  const double minE = e->fResMacXSecData[iStart + 5];
  if (ekin < minE) return 0.0;
  const double m = GetSplineLog3(nData, &e->fResMacXSecData[iStart + 5], ekin, lekin, e->fResMacXSecData[iStart + 3], e->fResMacXSecData[iStart + 4]);
  return dmax(0.0, m);
}
static double GetMacXSecNuclear(const struct G4HepEmElectronData* e, int imat, double ekin, double lekin) { /* :534-541 */
  if (ekin < e->fENucEnergyGrid[0]) return 0.0;
  const int n = e->fENucEnergyGridSize;
  const double m = GetSplineLog2(n, e->fENucEnergyGrid, &e->fENucMacXsecData[imat * 2 * n], ekin, lekin, e->fENucLogMinEkin, e->fENucEILDelta);
  return dmax(0.0, m);
}
static double GetRestMacXSecForStepping(const struct G4HepEmElectronData* e, int imc, double ekin, double lekin, int isioni) { /* :544-568 */
  const double log08 = -0.22314355131420971;
  const int iIoni = e->fResMacXSecStartIndexPerMatCut[imc];
  const int nIoni = (int)e->fResMacXSecData[iIoni];
  const int iStart = isioni ? iIoni : iIoni + 3 * nIoni + 5;
  const int nData = (int)e->fResMacXSecData[iStart];
  const double minE = e->fResMacXSecData[iStart + 5];
  const double maxE = e->fResMacXSecData[iStart + 1];
  const double maxV = e->fResMacXSecData[iStart + 2];
  if (ekin > maxE) {
    const double ekinReduced = 0.8 * ekin;
    if (ekinReduced < maxE) return dmax(0.0, maxV);
    ekin = ekinReduced;
    lekin += log08;
  }
  if (ekin < minE) return 0.0;
  const double m = GetSplineLog3(nData, &e->fResMacXSecData[iStart + 5], ekin, lekin, e->fResMacXSecData[iStart + 3], e->fResMacXSecData[iStart + 4]);
  return dmax(0.0, m);
}
static double GetTransportMFP(const struct G4HepEmElectronData* e, int im, double ekin, double lekin) { /* :576-583 */
  const int N = e->fELossEnergyGridSize;
  const double tr = dmax(0.0, GetSplineLog2(N, e->fELossEnergyGrid, &e->fTr1MacXSecData[2 * N * im], ekin, lekin, e->fELossLogMinEkin, e->fELossEILDelta));
  return tr > 0. ? 1. / tr : kALargeValue;
}
static double ComputeMacXsecAnnihilation(double ekin, double eDens) { /* :585-593 */
  const double tau = ekin * kInvElectronMassC2;
  const double gam = tau + 1.0;
  const double gam2 = gam * gam;
  const double bg2 = tau * (tau + 2.0);
  const double bg = sqrt(bg2);
  return eDens * kPir02 * ((gam2 + 4. * gam + 1.) * HLog(gam + bg) - (gam + 3.) * bg) / (bg2 * (gam + 1.));
}
static double ComputeMacXsecAnnihilationForStepping(double ekin, double eDens) { return ComputeMacXsecAnnihilation(0.8 * ekin, eDens); }

/* ---- G4HepEmElectronInteractionUMSC.icc -------------------------------------- */
static void UMSC_StepLimit(const struct G4HepEmData* d, const struct G4HepEmParameters* p, omsc* m, double ekin, int imat,
                           int ireg, double range, double presafety, int onBoundary, int isEl, orng* rng) { /* :19-126 */
  m->no_scat = 0;
  m->is_displace = 1;
  const double kTLimitMinfix = 1.0E-8;
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[imat];
  if (m->true_step < kTLimitMinfix || range * md->fUMSCPar < presafety) {
    m->is_displace = 0;
    return;
  }
  const struct G4HepEmRegionParmeters* rp = &p->fParametersPerRegion[ireg];
  const double rangeFactor = rp->fMSCRangeFactor, safetyFactor = rp->fMSCSafetyFactor;
  const int useSafety = !rp->fIsMSCMinimalStepLimit;
  double tlimit = 0.0;
  if (useSafety) {
    if (m->first_step || onBoundary) {
      const double lambdaTr1 = m->lambtr1;
      m->init_range = dmax(range, lambdaTr1);
      m->dyn_range_factor = lambdaTr1 > 1.0 ? rangeFactor * (0.75 + 0.25 * lambdaTr1) : rangeFactor;
      const double stepMin = lambdaTr1 * 1.0E-3 / (2.0E-3 + ekin * (md->fUMSCStepMinPars[0] + ekin * md->fUMSCStepMinPars[1]));
      const double dum0 = isEl ? 0.87 * md->fZeff23 : 0.70 * md->fZeffSqrt;
      const double dum1 = ekin > 5.0E-3 ? dum0 * stepMin : dum0 * stepMin * 0.5 * (1.0 + ekin * 200.0);
      m->tlimit_min = dmax(dum1, kTLimitMinfix);
      m->first_step = 0;
    }
    const double tlimitmin = m->tlimit_min;
    tlimit = range > presafety ? dmax(dmax(m->init_range * m->dyn_range_factor, safetyFactor * presafety), tlimitmin)
                               : dmax(range, tlimitmin);
  } else {
    if (onBoundary) {
      const double lambdaTr1 = m->lambtr1;
      const double tmp = range > lambdaTr1 ? rangeFactor * range : rangeFactor * lambdaTr1;
      m->init_range = dmax(tmp, 10 * kTLimitMinfix);
    }
    tlimit = m->init_range;
  }
  const double tlimitmin = m->tlimit_min;
  if (tlimit < m->true_step) {
    const double dum0 = tlimit > tlimitmin ? dmax(Gauss(rng, tlimit, 0.1 * (tlimit - tlimitmin)), tlimitmin) : tlimitmin;
    m->true_step = dmin(dum0, m->true_step);
  }
}

static double UMSC_SimpleScattering(double xmeanth, double x2meanth, orng* rng) { /* :280-293 */
  const double dum0 = 3. * x2meanth - 1;
  const double dum1 = 2. * xmeanth - dum0;
  const double a = 1. + 4. * dum0 / dum1;
  const double prob = (2. + a) * xmeanth / a;
  double r[2];
  flatArray(rng, 2, r);
  return (r[0] < prob) ? -1. + 2. * HPow(r[1], 1. / (1. + a)) : -1. + 2. * r[1];
}

static double UMSC_Theta0PositronCorrection(double eekin, double zeff) { /* :313-338 */
  const double ff = 1. + zeff * (1.84035E-4 * zeff - 1.86427E-2) + 0.41125;
  const double a = 0.994 - 4.08E-3 * zeff;
  const double b = 7.16 + (52.6 + 365. / zeff) / zeff;
  const double tu = sqrt(eekin) * kInvElectronMassC2;
  const double x = sqrt(tu * (tu + 2.) / ((tu + 1.) * (tu + 1.)));
  const double xl = 0.6;
  if (x < xl) return ff * a * (1. - HExp(-b * x));
  const double c = 1.00 - 4.47E-3 * zeff;
  const double dd = 1.21E-3 * zeff;
  const double e = 113.0;
  const double xh = 0.9;
  if (x > xh) return ff * (c + dd * HExp(e * (x - 1.)));
  const double yl = a * (1. - HExp(-b * xl));
  const double yh = c + dd * HExp(e * (xh - 1.));
  const double y0 = (yh - yl) / (xh - xl);
  const double y1 = yl - y0 * xl;
  return ff * (y0 * x + y1);
}

static double UMSC_ComputeTheta0(double stepInRadLength, double postEkin, double preEkin, double zeff, const double* thetaCoeff,
                                 int isEl, int isPosCor) { /* :297-309 */
  const double kHighland = 13.6;
  const double postInvBetaPc = (postEkin + kElectronMassC2) / (postEkin * (postEkin + 2. * kElectronMassC2));
  const double invBetaPc = preEkin != postEkin
      ? sqrt(postInvBetaPc * (preEkin + kElectronMassC2) / (preEkin * (preEkin + 2. * kElectronMassC2)))
      : postInvBetaPc;
  const double y = (isEl || !isPosCor) ? stepInRadLength : stepInRadLength * UMSC_Theta0PositronCorrection(preEkin * postEkin, zeff);
  return kHighland * sqrt(y) * invBetaPc * (thetaCoeff[0] + thetaCoeff[1] * HLog(y));
}

static double UMSC_SampleCosineTheta(double pStep, double preEkin, double preTr1mfp, double postEkin, double postTr1mfp,
                                     double tlimitMin, double radLength, double zeff, const double* tailCoeff,
                                     const double* thetaCoeff, int isEl, int isPosCor, orng* rng) { /* :152-277 */
  const double iPreTr1mfp = 1.0 / preTr1mfp;
  const double deltaR1mfp = preTr1mfp - postTr1mfp;
  const double tau = fabs(deltaR1mfp) > 0.01 * preTr1mfp ? pStep * HLog(preTr1mfp / postTr1mfp) / deltaR1mfp : pStep * iPreTr1mfp;
  const double kTauBig = 8.0;
  if (tau > kTauBig) return 2.0 * flat(rng) - 1.0;
  const double kTauSmall = 1.0E-16;
  if (tau < kTauSmall) return 1.0;
  double xmeanth, x2meanth;
  if (tau < 0.01) {
    xmeanth = 1.0 - tau * (1.0 - 0.5 * tau);
    x2meanth = 1.0 - tau * (5.0 - 6.25 * tau) * 0.333333;
  } else {
    xmeanth = HExp(-tau);
    x2meanth = (1.0 + 2.0 * HExp(-2.5 * tau)) * 0.333333;
  }
  if (postEkin < 0.5 * preEkin) return UMSC_SimpleScattering(xmeanth, x2meanth, rng);
  const double tsmall = dmin(tlimitMin, 1.0);
  const int notSmall = pStep > tsmall;
  const double theta0 = notSmall
      ? UMSC_ComputeTheta0(pStep / radLength, postEkin, preEkin, zeff, thetaCoeff, isEl, isPosCor)
      : UMSC_ComputeTheta0(tsmall / radLength, postEkin, preEkin, zeff, thetaCoeff, isEl, isPosCor) * sqrt(pStep / tsmall);
  if (theta0 > kPi * 0.166666) return UMSC_SimpleScattering(xmeanth, x2meanth, rng);
  const double theta2 = theta0 * theta0;
  if (theta2 < kTauSmall) return 1.0;
  const double dumtau = notSmall ? tau : tsmall * iPreTr1mfp;
  const double parU = HPow(dumtau, 0.1666666);
  const double dumxsi = tailCoeff[0] + parU * (tailCoeff[1] + parU * tailCoeff[2]) + tailCoeff[3] * HLog(pStep / (tau * radLength));
  const double parXsi = dmax(dumxsi, 1.9);
  const double parC = fabs(parXsi - 3.) < 0.001 ? 3.001 : fabs(parXsi - 2.) < 0.001 ? 2.001 : parXsi;
  const double dumC1 = parC - 1.;
  const double dumEa = HExp(-parXsi);
  const double dumEaa = 1. / (1. - dumEa);
  double thex = theta2 * (1.0 - theta2 * 0.0833333);
  if (theta2 > 0.01) {
    const double dum = 2.0 * sin(0.5 * theta0);
    thex = dum * dum;
  }
  const double xmean1 = 1. - (1. - (1. + parXsi) * dumEa) * thex * dumEaa;
  if (xmean1 <= 0.999 * xmeanth) return UMSC_SimpleScattering(xmeanth, x2meanth, rng);
  const double x0 = 1. - parXsi * thex;
  const double bx = parC * thex;
  const double b = bx + x0;
  const double b1 = b + 1.;
  const double eb1 = HPow(b1, dumC1);
  const double ebx = HPow(bx, dumC1);
  const double dd = ebx / eb1;
  const double xmean2 = (x0 + dd - (bx - b1 * dd) / (parC - 2.)) / (1. - dd);
  const double f1x0 = dumEa * dumEaa;
  const double f2x0 = dumC1 / (parC * (1. - dd));
  const double prob = f2x0 / (f1x0 + f2x0);
  const double qprb = xmeanth / (prob * xmean1 + (1. - prob) * xmean2);
  double r[3];
  flatArray(rng, 3, r);
  if (r[0] < qprb) {
    if (r[1] < prob) return 1. + HLog(dumEa + r[2] / dumEaa) * thex;
    const double var0 = (1.0 - dd) * r[2];
    if (var0 < 0.01 * dd) {
      const double var = var0 / (dd * dumC1);
      return -1.0 + var * (1.0 - var * 0.5 * parC) * b1;
    }
    return 1.0 + thex * (parC - parXsi - parC * HPow(var0 + dd, -1. / dumC1));
  }
  return 2.0 * r[1] - 1.0;
}

static void UMSC_SampleDisplacement(double pStep, double thePhi, omsc* m, orng* rng) { /* :342-357 */
  const double r = 0.73 * sqrt((pStep - m->zpath) * (pStep + m->zpath));
  const double cbeta = 2.16;
  const double cbeta1 = 1. - HExp(-cbeta * kPi);
  double rr[2];
  flatArray(rng, 2, rr);
  const double psi = -HLog(1. - rr[0] * cbeta1) / cbeta;
  const double phi = (rr[1] < 0.5) ? thePhi + psi : thePhi - psi;
  m->displ[0] = r * cos(phi);
  m->displ[1] = r * sin(phi);
  m->displ[2] = 0.0;
}

static void UMSC_SampleScattering(const struct G4HepEmData* d, omsc* m, double pStep, double preEkin, double preTr1mfp,
                                  double postEkin, double postTr1mfp, int imat, int isEl, int isPosCor, int isDisplace,
                                  orng* rng) { /* :129-149 */
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[imat];
  const double cost = UMSC_SampleCosineTheta(pStep, preEkin, preTr1mfp, postEkin, postTr1mfp, m->tlimit_min, md->fRadiationLength,
                                             md->fZeff, md->fUMSCTailCoeff, md->fUMSCThetaCoeff, isEl, isPosCor, rng);
  if (fabs(cost) >= 1.0) {
    m->no_scat = 1;
    return;
  }
  const double sth = sqrt((1.0 - cost) * (1.0 + cost));
  const double phi = k2Pi * flat(rng);
  m->dir[0] = sth * cos(phi);
  m->dir[1] = sth * sin(phi);
  m->dir[2] = cost;
  m->is_displace = m->is_displace && isDisplace;
  if (m->is_displace && pStep > m->zpath) UMSC_SampleDisplacement(pStep, phi, m, rng);
}

/* ---- G4HepEmElectronEnergyLossFluctuation.icc -------------------------------- */
static double SampleGaussianLoss(double meane, double sig2e, orng* rng) { /* :100-111 */
  const double twom = 2. * meane;
  if (meane * meane < 0.0625 * sig2e) return twom * flat(rng);
  const double sig = sqrt(sig2e);
  double eloss;
  do {
    eloss = Gauss(rng, meane, sig);
  } while (eloss < 0. || eloss > twom);
  return eloss;
}

static double SampleEnergyLossFluctuation(double tcut, double excEner, double meanELoss, orng* rng) { /* :11-97 */
  const double scaling = dmin(1. + 5.E-4 / tcut, 1.5);
  const double meanLoss = meanELoss / scaling;
  const double kRate = 0.56, kE0 = 1.E-5, kNMaxCont = 8.;
  const double w1 = tcut / kE0;
  double a3 = meanLoss * (tcut - kE0) / (kE0 * tcut * HLog(w1));
  double a1 = 0., e1 = excEner, eloss = 0.0;
  if (tcut > excEner) {
    const double a1Tmp = meanLoss * (1. - kRate) / excEner;
    const double kA0 = 42., kFw = 4.;
    const double dum0 = a1Tmp < kA0 ? .1 + (kFw - .1) * sqrt(a1Tmp / kA0) : kFw;
    a1 = a1Tmp / dum0;
    e1 *= dum0;
    a3 *= kRate;
    if (a1 > kNMaxCont) {
      const double emean = a1 * e1;
      const double sig2e = emean * e1;
      eloss = SampleGaussianLoss(emean, sig2e, rng);
    } else {
      const int pp = Poisson(rng, a1);
      eloss = pp > 0 ? ((pp + 1) - 2. * flat(rng)) * e1 : 0.;
    }
  }
  if (a3 > 0.) {
    double p3 = a3, alfa = 1.;
    if (a3 > kNMaxCont) {
      alfa = w1 * (kNMaxCont + a3) / (w1 * kNMaxCont + a3);
      const double alfa1 = alfa * HLog(alfa) / (alfa - 1.);
      const double namean = a3 * w1 * (alfa - 1.) / ((w1 - 1.) * alfa);
      const double emean = namean * kE0 * alfa1;
      const double sig2e = kE0 * kE0 * namean * (alfa - alfa1 * alfa1);
      eloss += SampleGaussianLoss(emean, sig2e, rng);
      p3 = a3 - namean;
    }
    const double w3 = alfa * kE0;
    if (tcut > w3) {
      const double w = (tcut - w3) / tcut;
      const int nnb = Poisson(rng, p3);
      if (nnb > 0) {
        const int kBlock = 8;
        const int nBlocks = nnb / kBlock;
        double rndm[8];
        for (int ib = 0; ib < nBlocks; ++ib) {
          flatArray(rng, kBlock, rndm);
          for (int i = 0; i < kBlock; ++i) eloss += w3 / (1. - w * rndm[i]);
        }
        const int nTail = nnb - nBlocks * kBlock;
        flatArray(rng, nTail, rndm);
        for (int i = 0; i < nTail; ++i) eloss += w3 / (1. - w * rndm[i]);
      }
    }
  }
  return eloss * scaling;
}

/* ---- G4HepEmInteractionUtils.{hh,icc} ----------------------------------------- */
/* LPM G(s), Phi(s) table on s in [0,2] step 0.05 (Stanev et al. approximation
   tabulated in G4HepEmInteractionUtils.hh:21-36 and in Geant4's G4eBremsstrahlungRelModel) */
static const double kFuncLPM[] = {
  0.0000E+00, 0.0000E+00, 6.9163E-02, 2.5747E-01, 2.0597E-01, 4.4573E-01,
  3.5098E-01, 5.8373E-01, 4.8095E-01, 6.8530E-01, 5.8926E-01, 7.6040E-01,
  6.7626E-01, 8.1626E-01, 7.4479E-01, 8.5805E-01, 7.9826E-01, 8.8952E-01,
  8.4003E-01, 9.1338E-01, 8.7258E-01, 9.3159E-01, 8.9794E-01, 9.4558E-01,
  9.1776E-01, 9.5640E-01, 9.3332E-01, 9.6483E-01, 9.4560E-01, 9.7143E-01,
  9.5535E-01, 9.7664E-01, 9.6313E-01, 9.8078E-01, 9.6939E-01, 9.8408E-01,
  9.7444E-01, 9.8673E-01, 9.7855E-01, 9.8888E-01, 9.8191E-01, 9.9062E-01,
  9.8467E-01, 9.9204E-01, 9.8695E-01, 9.9321E-01, 9.8884E-01, 9.9417E-01,
  9.9042E-01, 9.9497E-01, 9.9174E-01, 9.9564E-01, 9.9285E-01, 9.9619E-01,
  9.9379E-01, 9.9666E-01, 9.9458E-01, 9.9706E-01, 9.9526E-01, 9.9739E-01,
  9.9583E-01, 9.9768E-01, 9.9632E-01, 9.9794E-01, 9.9674E-01, 9.9818E-01,
  9.9710E-01, 9.9839E-01, 9.9741E-01, 9.9857E-01, 9.9767E-01, 9.9873E-01,
  9.9790E-01, 9.9887E-01, 9.9809E-01, 9.9898E-01, 9.9826E-01, 9.9909E-01,
  9.9840E-01, 9.9918E-01, 9.9856E-01, 9.9926E-01};

static double SampleCostModifiedTsai(double primEkin, orng* rng) { /* .icc:11-25 */
  const double uMax = 2.0 * (1.0 + primEkin * kInvElectronMassC2);
  double r[3], u;
  do {
    flatArray(rng, 3, r);
    const double uu = -HLog(r[0] * r[1]);
    u = (0.25 > r[2]) ? uu * 1.6 : uu * 0.533333333;
  } while (u > uMax);
  return 1.0 - 2.0 * u * u / (uMax * uMax);
}

static void EvaluateLPMFunctions(double* funcXiS, double* funcGS, double* funcPhiS, double egamma, double etotal, double elpm,
                                 double z23, double ilVarS1, double ilVarS1Cond, double densityCor, double times) { /* .icc:28-76 */
  const double sqrt2 = 1.414213562373095;
  const double redegamma = egamma / etotal;
  const double varSprime = sqrt(0.125 * redegamma * elpm / (times * (1.0 - redegamma) * etotal));
  const double varS1 = z23 / (184.15 * 184.15);
  const double condition = sqrt2 * varS1;
  double funcXiSprime = 2.0;
  if (varSprime > 1.0) {
    funcXiSprime = 1.0;
  } else if (varSprime > condition) {
    const double funcHSprime = HLog(varSprime) * ilVarS1Cond;
    funcXiSprime = 1.0 + funcHSprime - 0.08 * (1.0 - funcHSprime) * funcHSprime * (2.0 - funcHSprime) * ilVarS1Cond;
  }
  *funcXiS = funcXiSprime;
  const double varS = varSprime / sqrt(funcXiSprime);
  double varShat = varS;
  if (densityCor != 0.0) {
    varShat *= (1.0 + densityCor / (egamma * egamma));
    *funcXiS = 2.0;
    if (varShat > 1.0) {
      *funcXiS = 1.0;
    } else if (varShat > varS1) {
      *funcXiS = 1.0 + HLog(varShat) * ilVarS1;
    }
  }
  const double lpmSLimit = 2.0, lpmISDelt = 20.0;
  if (varShat < lpmSLimit) {
    double val = varShat * lpmISDelt;
    int ilow = (int)val;
    val -= ilow;
    ilow *= 2;
    *funcGS = (kFuncLPM[ilow + 2] - kFuncLPM[ilow]) * val + kFuncLPM[ilow];
    *funcPhiS = (kFuncLPM[ilow + 3] - kFuncLPM[ilow + 1]) * val + kFuncLPM[ilow + 1];
  } else {
    double ss = 1.0 / (varShat * varShat);
    ss *= ss;
    *funcGS = 1.0 - 0.0230655 * ss;
    *funcPhiS = 1.0 - 0.01190476 * ss;
  }
  if ((*funcXiS) * (*funcPhiS) > 1.0 || varShat > 0.57) *funcXiS = 1.0 / (*funcPhiS);
}

/* ---- G4HepEmElectronInteractionIoni.icc ------------------------------------------ */
static double SampleETransferMoller(double elCut, double primEkin, orng* rng) { /* :50-74 */
  const double tmin = elCut, tmax = 0.5 * primEkin;
  const double xmin = tmin / primEkin, xmax = tmax / primEkin;
  const double gamma = primEkin * kInvElectronMassC2 + 1.0;
  const double gamma2 = gamma * gamma;
  const double xminmax = xmin * xmax;
  const double gg = (2.0 * gamma - 1.0) / gamma2;
  const double y = 1. - xmax;
  const double gf = 1.0 - gg * xmax + xmax * xmax * (1.0 - gg + (1.0 - gg * y) / (y * y));
  double dum, r[2], delta = 0.;
  do {
    flatArray(rng, 2, r);
    delta = xminmax / (xmin * (1.0 - r[0]) + xmax * r[0]);
    const double xx = 1.0 - delta;
    dum = 1.0 - gg * delta + delta * delta * (1.0 - gg + (1.0 - gg * xx) / (xx * xx));
  } while (gf * r[1] > dum);
  return delta * primEkin;
}
static double SampleETransferBhabha(double elCut, double primEkin, orng* rng) { /* :76-108 */
  const double tmin = elCut, tmax = primEkin;
  const double xmin = tmin / primEkin, xmax = tmax / primEkin;
  const double gamma = primEkin * kInvElectronMassC2 + 1.0;
  const double gamma2 = gamma * gamma;
  const double beta2 = 1. - 1. / gamma2;
  const double xminmax = xmin * xmax;
  const double y = 1.0 / (1.0 + gamma), y2 = y * y, y12 = 1.0 - 2.0 * y;
  const double b1 = 2.0 - y2, b2 = y12 * (3.0 + y2), y122 = y12 * y12, b4 = y122 * y12, b3 = b4 + y122;
  const double xmax2 = xmax * xmax;
  const double gf = 1.0 + (xmax2 * b4 - xmin * xmin * xmin * b3 + xmax2 * b2 - xmin * b1) * beta2;
  double dum, r[2], delta = 0.;
  do {
    flatArray(rng, 2, r);
    delta = xminmax / (xmin * (1.0 - r[0]) + xmax * r[0]);
    const double xx = delta * delta;
    dum = 1.0 + (xx * xx * b4 - delta * xx * b3 + xx * b2 - delta * b1) * beta2;
  } while (gf * r[1] > dum);
  return delta * primEkin;
}
static void IoniSampleDirections(double primEkin, double deltaEkin, double* secDir, double* primDir, orng* rng) { /* :111-138 */
  const double elInitETot = primEkin + kElectronMassC2;
  const double elInitPTot = sqrt(primEkin * (elInitETot + kElectronMassC2));
  const double deltaPTot = sqrt(deltaEkin * (deltaEkin + 2.0 * kElectronMassC2));
  const double cost = deltaEkin * (elInitETot + kElectronMassC2) / (deltaPTot * elInitPTot);
  const double cosTheta = dmax(-1.0, dmin(cost, 1.0));
  const double sinTheta = sqrt((1.0 - cosTheta) * (1.0 + cosTheta));
  const double phi = k2Pi * flat(rng);
  secDir[0] = sinTheta * cos(phi);
  secDir[1] = sinTheta * sin(phi);
  secDir[2] = cosTheta;
  RotateToReferenceFrame(secDir, primDir);
  primDir[0] = elInitPTot * primDir[0] - deltaPTot * secDir[0];
  primDir[1] = elInitPTot * primDir[1] - deltaPTot * secDir[1];
  primDir[2] = elInitPTot * primDir[2] - deltaPTot * secDir[2];
  const double norm = 1.0 / sqrt(primDir[0] * primDir[0] + primDir[1] * primDir[1] + primDir[2] * primDir[2]);
  primDir[0] *= norm;
  primDir[1] *= norm;
  primDir[2] *= norm;
}
static void IoniPerform(otl* tl, const struct G4HepEmData* d, int isEl) { /* :19-47 */
  otrack* t = &tl->el.t;
  const double primEkin = t->ekin;
  const double elCut = d->fTheMatCutData->fMatCutData[t->imc].fSecElProdCutE;
  const double maxET = isEl ? 0.5 * primEkin : primEkin;
  if (maxET <= elCut) return;
  const double delta = isEl ? SampleETransferMoller(elCut, primEkin, tl->rng) : SampleETransferBhabha(elCut, primEkin, tl->rng);
  osec* s = AddSecEl(tl, HEPEM_ELECTRON);
  IoniSampleDirections(primEkin, delta, s->dir, t->dir, tl->rng);
  SetEKin(t, primEkin - delta);
  s->ekin = delta;
}

/* ---- G4HepEmElectronInteractionBrem.icc ------------------------------------------ */
static int BremSelectTargetAtom(const struct G4HepEmElectronData* e, int imc, double ekin, double lekin, double urndn,
                                int isSB) { /* :266-297 */
  const int start = isSB ? e->fElemSelectorBremSBStartIndexPerMatCut[imc] : e->fElemSelectorBremRBStartIndexPerMatCut[imc];
  const double* data = isSB ? &e->fElemSelectorBremSBData[start] : &e->fElemSelectorBremRBData[start];
  const int numData = (int)data[0];
  const int numElem = (int)data[1];
  const double logE0 = data[2], invLD = data[3];
  const double* xdata = &data[4];
  const double xv = dmax(xdata[0], dmin(xdata[numElem * (numData - 1)], ekin));
  const int idxEkin = (int)dmax(0.0, dmin((lekin - logE0) * invLD, numData - 2.0));
  int i0 = idxEkin * numElem, i1 = i0 + numElem;
  const double x1 = xdata[i0++], x2 = xdata[i1++];
  const double dl = x2 - x1;
  const double b = dmax(0., dmin(1., (xv - x1) / dl));
  int ie = 0;
  while (ie < numElem - 1 && urndn > xdata[i0 + ie] + b * (xdata[i1 + ie] - xdata[i0 + ie])) ++ie;
  return ie;
}
static int BremLinSearch(const double* v, int size, double val) { /* :328-344 */
  int i = 0;
  const int size3 = 3 * size;
  while (i + 9 < size3) {
    if (v[i + 0] > val) return i + 0;
    if (v[i + 3] > val) return i + 3;
    if (v[i + 6] > val) return i + 6;
    if (v[i + 9] > val) return i + 9;
    i += 12;
  }
  while (i < size3) {
    if (v[i] > val) break;
    i += 3;
  }
  return i;
}
static double SampleETransferSB(const struct G4HepEmData* d, double primEkin, double lEkin, int imc, orng* rng, int isEl) { /* :73-182 */
  const struct G4HepEmMCCData* mc = &d->fTheMatCutData->fMatCutData[imc];
  const double gamCut = mc->fSecGamProdCutE, logGamCut = mc->fLogSecGamCutE;
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[mc->fHepEmMatIndex];
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const int elemIndx = (md->fNumOfElement > 1) ? BremSelectTargetAtom(e, imc, primEkin, lEkin, flat(rng), 1) : 0;
  const int iZet = md->fElementVect[elemIndx];
  const double dZet = (double)iZet;
  const struct G4HepEmSBTableData* sb = d->fTheSBTableData;
  const int iStart = sb->fSBTablesStartPerZ[iZet];
  const int iGamCut = sb->fGammaCutIndices[sb->fGammaCutIndxStartIndexPerMC[imc] + elemIndx];
  int isCorner = 0, isSimply = 0;
  int eIdx = (int)(sb->fSBTableData[iStart + 2]);
  if (primEkin < sb->fElEnergyVect[eIdx]) {
    const double val = (lEkin - sb->fLogMinElEnergy) * sb->fILDeltaElEnergy;
    eIdx = (int)val;
    double pIndxH = val - eIdx;
    if (sb->fElEnergyVect[eIdx] <= gamCut) {
      pIndxH = (lEkin - logGamCut) / (sb->fLElEnergyVect[eIdx + 1] - logGamCut);
      isCorner = 1;
    }
    if (flat(rng) < pIndxH) ++eIdx;
    else if (isCorner) isSimply = 1;
  }
  const int minEIndx = (int)(sb->fSBTableData[iStart + 1]);
  const int numGamCuts = (int)(sb->fSBTableData[iStart + 3]);
  const int sizeOneE = (int)(numGamCuts + 3 * sb->fNumKappa);
  const int iSTStart = iStart + 4 + (eIdx - minEIndx) * sizeOneE;
  const double minV = sb->fSBTableData[iSTStart + iGamCut];
  const double* stData = &sb->fSBTableData[iSTStart + numGamCuts];
  const double lKTrans = (logGamCut - lEkin) / (logGamCut - sb->fLElEnergyVect[eIdx]);
  const double primETot = primEkin + kElectronMassC2;
  const double dielSupConst = md->fDensityCorFactor * primETot * primETot;
  double suppression = 1.0, r[2], eGamma = 0.0;
  do {
    flatArray(rng, 2, r);
    double kappa = 1.0;
    if (!isSimply) {
      const double cumRV = r[0] * (1.0 - minV) + minV;
      const int cumLIndx3 = BremLinSearch(stData, sb->fNumKappa, cumRV) - 3;
      const int cumLIndx = cumLIndx3 / 3;
      const double cumL = stData[cumLIndx3];
      const double pA = stData[cumLIndx3 + 1];
      const double pB = stData[cumLIndx3 + 2];
      const double cumH = stData[cumLIndx3 + 3];
      const double lKL = sb->fLKappaVect[cumLIndx];
      const double lKH = sb->fLKappaVect[cumLIndx + 1];
      const double dm1 = (cumRV - cumL) / (cumH - cumL);
      const double dm2 = (1.0 + pA + pB) * dm1;
      const double dm3 = 1.0 + dm1 * (pA + pB * dm1);
      const double lKappa = lKL + dm2 / dm3 * (lKH - lKL);
      kappa = HExp(lKappa * lKTrans);
    } else {
      kappa = 1.0 - r[0] * (1.0 - gamCut / primEkin);
    }
    eGamma = kappa * primEkin;
    const double invEGamma = 1.0 / eGamma;
    suppression = 1.0 / (1.0 + dielSupConst * invEGamma * invEGamma);
    if (!isEl) {
      const double e1 = primEkin - gamCut;
      const double iBeta1 = (e1 + kElectronMassC2) / sqrt(e1 * (e1 + 2.0 * kElectronMassC2));
      const double e2 = primEkin - eGamma;
      const double iBeta2 = (e2 + kElectronMassC2) / sqrt(e2 * (e2 + 2.0 * kElectronMassC2));
      const double dum = kAlpha * k2Pi * dZet * (iBeta1 - iBeta2);
      suppression = (dum > -12.) ? suppression * HExp(dum) : 0.;
    }
  } while (r[1] > suppression);
  return eGamma;
}
static double SampleETransferRB(const struct G4HepEmData* d, double primEkin, double lEkin, int imc, orng* rng, int isEl) { /* :184-263 */
  const struct G4HepEmMCCData* mc = &d->fTheMatCutData->fMatCutData[imc];
  const double gamCut = mc->fSecGamProdCutE;
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[mc->fHepEmMatIndex];
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const int elemIndx = (md->fNumOfElement > 1) ? BremSelectTargetAtom(e, imc, primEkin, lEkin, flat(rng), 0) : 0;
  const int iZet = md->fElementVect[elemIndx];
  const double dZet = (double)iZet;
  const int zcl = iZet < d->fTheElementData->fMaxZet ? iZet : d->fTheElementData->fMaxZet;
  const struct G4HepEmElemData* ed = &d->fTheElementData->fElementData[zcl];
  const double densityFactor = kMigdalConst * md->fElectronDensity;
  const double lpmEnergy = kLPMconstant * md->fRadiationLength;
  const double lpmEnergyLim = sqrt(densityFactor) * lpmEnergy;
  const double primTotalE = primEkin + kElectronMassC2;
  const double densityCorr = densityFactor * primTotalE * primTotalE;
  const int isLPMActive = (primTotalE > lpmEnergyLim);
  const double zFactor1 = ed->fZFactor1;
  const double zFactor2 = (1. + 1. / dZet) / 12.;
  const double rejFuncMax = zFactor1 + zFactor2;
  const double xmin = HLog(gamCut * gamCut + densityCorr);
  const double xrange = HLog(primEkin * primEkin + densityCorr) - xmin;
  double r[2], eGamma, funcVal;
  do {
    flatArray(rng, 2, r);
    eGamma = sqrt(dmax(HExp(xmin + r[0] * xrange) - densityCorr, 0.0));
    const double y = eGamma / primTotalE;
    const double onemy = 1. - y;
    const double dum0 = 0.25 * y * y;
    if (isLPMActive) {
      double funcGS, funcPhiS, funcXiS;
      EvaluateLPMFunctions(&funcXiS, &funcGS, &funcPhiS, eGamma, primTotalE, lpmEnergy, ed->fZet23, ed->fILVarS1, ed->fILVarS1Cond,
                           densityCorr, 1.0);
      const double term1 = funcXiS * (dum0 * funcGS + (onemy + 2.0 * dum0) * funcPhiS);
      funcVal = term1 * zFactor1 + onemy * zFactor2;
    } else {
      const double dum1 = onemy + 3. * dum0;
      if (iZet < 5) {
        funcVal = dum1 * zFactor1 + onemy * zFactor2;
      } else {
        const double dum2 = y / (primTotalE - eGamma);
        const double gam = dum2 * 100. * kElectronMassC2 / ed->fZet13;
        const double eps = gam / ed->fZet13;
        const double gam2 = gam * gam;
        const double phi1 = 16.863 - 2.0 * HLog(1.0 + 0.311877 * gam2) + 2.4 * HExp(-0.9 * gam) + 1.6 * HExp(-1.5 * gam);
        const double phi2 = 2.0 / (3.0 + 19.5 * gam + 18.0 * gam2);
        const double eps2 = eps * eps;
        const double psi1 = 24.34 - 2.0 * HLog(1.0 + 13.111641 * eps2) + 2.8 * HExp(-8.0 * eps) + 1.2 * HExp(-29.2 * eps);
        const double psi2 = 2.0 / (3.0 + 120.0 * eps + 1200.0 * eps2);
        const double logZ = ed->fLogZ;
        const double Fz = logZ / 3. + ed->fCoulomb;
        const double invZ = 1. / dZet;
        funcVal = dum1 * ((0.25 * phi1 - Fz) + (0.25 * psi1 - 2. * logZ / 3.) * invZ) + 0.125 * onemy * (phi2 + psi2 * invZ);
      }
    }
    funcVal = dmax(0.0, funcVal);
  } while (funcVal < rejFuncMax * r[1]);
  return eGamma;
}
static void BremSampleDirections(double primEkin, double secEkin, double* secDir, double* primDir, orng* rng) { /* :299-322 */
  const double cost = SampleCostModifiedTsai(primEkin, rng);
  const double sint = sqrt((1.0 - cost) * (1.0 + cost));
  const double phi = k2Pi * flat(rng);
  secDir[0] = sint * cos(phi);
  secDir[1] = sint * sin(phi);
  secDir[2] = cost;
  RotateToReferenceFrame(secDir, primDir);
  const double primETot = primEkin + kElectronMassC2;
  const double primPTot = sqrt(primEkin * (primETot + kElectronMassC2));
  primDir[0] = primPTot * primDir[0] - secEkin * secDir[0];
  primDir[1] = primPTot * primDir[1] - secEkin * secDir[1];
  primDir[2] = primPTot * primDir[2] - secEkin * secDir[2];
  const double norm = 1.0 / sqrt(primDir[0] * primDir[0] + primDir[1] * primDir[1] + primDir[2] * primDir[2]);
  primDir[0] *= norm;
  primDir[1] *= norm;
  primDir[2] *= norm;
}
static void BremPerform(otl* tl, const struct G4HepEmData* d, int isEl, int isSB) { /* :35-70 */
  otrack* t = &tl->el.t;
  const double primEkin = t->ekin;
  const double lEkin = GetLogEKin(t);
  const double gamCut = d->fTheMatCutData->fMatCutData[t->imc].fSecGamProdCutE;
  if (primEkin <= gamCut) return;
  const double eGamma = isSB ? SampleETransferSB(d, primEkin, lEkin, t->imc, tl->rng, isEl)
                             : SampleETransferRB(d, primEkin, lEkin, t->imc, tl->rng, isEl);
  osec* s = AddSecGm(tl);
  BremSampleDirections(primEkin, eGamma, s->dir, t->dir, tl->rng);
  SetEKin(t, primEkin - eGamma);
  s->ekin = eGamma;
}

/* ---- G4HepEmPositronInteractionAnnihilation.icc ------------------------------- */
static void AnnihilateAtRest(otl* tl) { /* :23-49 */
  const double cost = 2. * flat(tl->rng) - 1.;
  const double sint = sqrt((1. - cost) * (1. + cost));
  const double phi = k2Pi * flat(tl->rng);
  osec* g1 = AddSecGm(tl);
  osec* g2 = AddSecGm(tl);
  g1->dir[0] = sint * cos(phi);
  g1->dir[1] = sint * sin(phi);
  g1->dir[2] = cost;
  g2->dir[0] = -g1->dir[0];
  g2->dir[1] = -g1->dir[1];
  g2->dir[2] = -g1->dir[2];
  g1->ekin = kElectronMassC2;
  g2->ekin = kElectronMassC2;
}
static void AnnihilateInFlight(otl* tl) { /* :51-119 */
  otrack* t = &tl->el.t;
  const double primEkin = t->ekin;
  const double* primDir = t->dir;
  osec* g1 = AddSecGm(tl);
  osec* g2 = AddSecGm(tl);
  const double tau = primEkin * kInvElectronMassC2;
  const double gam = tau + 1.0;
  const double tau2 = tau + 2.0;
  const double sqgrate = sqrt(tau / tau2) * 0.5;
  const double epsmin = 0.5 - sqgrate, epsmax = 0.5 + sqgrate;
  const double epsqot = epsmax / epsmin;
  const double tau4 = tau2 * tau2;
  double eps = 0.0, rfunc = 0.0, r[2];
  do {
    flatArray(tl->rng, 2, r);
    eps = epsmin * HExp(HLog(epsqot) * r[0]);
    rfunc = 1. - eps + (2. * gam * eps - 1.) / (eps * tau4);
  } while (rfunc < r[1]);
  const double sqg2m1 = sqrt(tau * tau2);
  const double cost = dmin(1., dmax(-1., (eps * tau2 - 1.) / (eps * sqg2m1)));
  const double sint = sqrt((1. + cost) * (1. - cost));
  const double phi = k2Pi * flat(tl->rng);
  const double initEt = primEkin + 2. * kElectronMassC2;
  const double ekinG1 = eps * initEt;
  g1->ekin = ekinG1;
  g1->dir[0] = sint * cos(phi);
  g1->dir[1] = sint * sin(phi);
  g1->dir[2] = cost;
  RotateToReferenceFrame(g1->dir, primDir);
  g2->ekin = initEt - ekinG1;
  const double initPt = sqrt(primEkin * (primEkin + 2 * kElectronMassC2));
  const double px = initPt * primDir[0] - g1->dir[0] * ekinG1;
  const double py = initPt * primDir[1] - g1->dir[1] * ekinG1;
  const double pz = initPt * primDir[2] - g1->dir[2] * ekinG1;
  const double norm = 1.0 / sqrt(px * px + py * py + pz * pz);
  g2->dir[0] = px * norm;
  g2->dir[1] = py * norm;
  g2->dir[2] = pz * norm;
  SetEKin(t, 0.0);
}

/* ---- G4HepEmElectronManager.icc: stepping ------------------------------------ */
static void ConvertTrueToGeometricLength(const struct G4HepEmData* d, omsc* m, double ekin, double range, int imc, int isEl) { /* :602-650 */
  m->par1 = -1.;
  m->par2 = 0.;
  m->par3 = 0.;
  m->true_step = dmin(m->true_step, range);
  m->zpath = m->true_step;
  const double kTlimitMinfix2 = 1.0E-6;
  if (m->true_step < kTlimitMinfix2) return;
  const double kTauSmall = 1.0e-16, kDtrl = 0.05;
  const double tau = m->true_step / m->lambtr1;
  if (tau < kTauSmall) {
    m->zpath = dmin(m->true_step, m->lambtr1);
  } else if (m->true_step < range * kDtrl) {
    const double kTauLim = 1.0e-6;
    m->zpath = (tau < kTauLim) ? m->true_step * (1. - 0.5 * tau) : m->lambtr1 * (1. - HExp(-tau));
  } else if (ekin < kElectronMassC2 || m->true_step == range) {
    m->par1 = 1. / range;
    m->par2 = 1. / (m->par1 * m->lambtr1);
    m->par3 = 1. + m->par2;
    m->zpath = 1. / (m->par1 * m->par3);
    if (m->true_step < range) m->zpath *= (1. - HPow(1. - m->true_step / range, m->par3));
  } else {
    const double rfin = dmax(range - m->true_step, 0.01 * range);
    const struct G4HepEmElectronData* e = ElData(d, isEl);
    const double t1 = GetInvRange(e, imc, rfin);
    const int imat = d->fTheMatCutData->fMatCutData[imc].fHepEmMatIndex;
    const double lambda1 = GetTransportMFP(e, imat, t1, HLog(t1));
    m->par1 = (m->lambtr1 - lambda1) / (m->lambtr1 * m->true_step);
    m->par2 = 1. / (m->par1 * m->lambtr1);
    m->par3 = 1. + m->par2;
    m->zpath = (1. - HPow(lambda1 / m->lambtr1, m->par3)) / (m->par1 * m->par3);
  }
  m->zpath = dmin(m->zpath, m->lambtr1);
}
static void ConvertGeometricToTrueLength(omsc* m, double range, double gStep) { /* :653-690 */
  m->zpath = gStep;
  const double kTLimitMinfix2 = 1.0E-6;
  if (gStep < kTLimitMinfix2) {
    m->true_step = gStep;
    return;
  }
  const double kTauSmall = 1.0e-16;
  double tlength = gStep;
  if (gStep > m->lambtr1 * kTauSmall) {
    if (m->par1 < 0.) {
      tlength = -m->lambtr1 * HLog(1. - gStep / m->lambtr1);
    } else {
      const double dum = m->par1 * m->par3 * gStep;
      tlength = (dum < 1.) ? (1. - HPow(1. - dum, 1. / m->par3)) / m->par1 : range;
    }
    if (tlength < gStep) tlength = gStep;
  }
  m->true_step = tlength;
}

static void HowFarToDiscreteInteraction(const struct G4HepEmData* d, const struct G4HepEmParameters* p, oel* el) { /* :48-101 */
  int winner = -1;
  otrack* t = &el->t;
  const double ekin = t->ekin;
  const double lekin = GetLogEKin(t);
  const int imc = t->imc;
  const int isEl = t->charge < 0.0;
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const double range = GetRestRange(e, imc, ekin, lekin);
  el->range = range;
  const int ireg = d->fTheMatCutData->fMatCutData[imc].fG4RegionIndex;
  const double frange = p->fParametersPerRegion[ireg].fFinalRange;
  const double drange = p->fParametersPerRegion[ireg].fDRoverRange;
  double pStep = (range > frange) ? range * drange + frange * (1.0 - drange) * (2.0 - frange / range) : range;
  const int imat = d->fTheMatCutData->fMatCutData[imc].fHepEmMatIndex;
  double mx[4];
  mx[0] = GetRestMacXSecForStepping(e, imc, ekin, lekin, 1);
  mx[1] = GetRestMacXSecForStepping(e, imc, ekin, lekin, 0);
  mx[2] = isEl ? 0.0 : ComputeMacXsecAnnihilationForStepping(ekin, d->fTheMaterialData->fMaterialData[imat].fElectronDensity);
  mx[3] = GetMacXSecNuclear(e, imat, ekin, lekin);
  for (int ip = 0; ip < 4; ++ip) {
    const double mfp = (mx[ip] > 0.) ? 1. / mx[ip] : kALargeValue;
    t->mfp[ip] = mfp;
    const double dStep = mfp * t->nial[ip];
    if (dStep < pStep) {
      pStep = dStep;
      winner = ip;
    }
  }
  el->pstep = pStep;
  t->winner = winner;
  t->gstep = pStep;
}

static void HowFarToMSC(const struct G4HepEmData* d, const struct G4HepEmParameters* p, oel* el, orng* rng) { /* :103-164 */
  double pStep = el->pstep;
  const double range = el->range;
  otrack* t = &el->t;
  const double ekin = t->ekin;
  const double lekin = GetLogEKin(t);
  const int imc = t->imc;
  const int isEl = t->charge < 0.0;
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const struct G4HepEmMCCData* mc = &d->fTheMatCutData->fMatCutData[imc];
  const int imat = mc->fHepEmMatIndex, ireg = mc->fG4RegionIndex;
  omsc* m = &el->msc;
  m->true_step = pStep;
  m->zpath = pStep;
  m->active = 0;
  m->displ[0] = m->displ[1] = m->displ[2] = 0.;
  m->dir[0] = 0.; m->dir[1] = 0.; m->dir[2] = 1.;
  const double kGeomMinLength = 5.E-8;
  if (pStep > kGeomMinLength && ekin > 1.0E-3) {
    m->active = 1;
    m->lambtr1 = GetTransportMFP(e, imat, ekin, lekin);
    UMSC_StepLimit(d, p, m, ekin, imat, ireg, range, t->safety, t->on_boundary, isEl, rng);
    ConvertTrueToGeometricLength(d, m, ekin, range, imc, isEl);
    const double mscTrue = m->true_step;
    if (mscTrue < pStep) {
      t->winner = -2;
      pStep = mscTrue;
      el->pstep = pStep;
    }
    t->gstep = dmin(m->zpath, pStep);
  }
}

static void UpdatePStepLength(oel* el) { /* :171-203 */
  otrack* t = &el->t;
  const double gStep = t->gstep;
  double pStep = gStep;
  const double range = el->range;
  omsc* m = &el->msc;
  if (m->active) {
    pStep = m->true_step;
    if (gStep < m->zpath) {
      ConvertGeometricToTrueLength(m, range, gStep);
      pStep = dmin(pStep, m->true_step);
      m->true_step = pStep;
    }
    const double kGeomMinLength = 5.E-8;
    if (pStep <= kGeomMinLength || range <= pStep) m->active = 0;
  }
  el->pstep = pStep;
}

static void UpdateNumIALeft(oel* el) { /* :205-214 */
  const double pStep = el->pstep;
  otrack* t = &el->t;
  t->nial[0] -= pStep / t->mfp[0];
  t->nial[1] -= pStep / t->mfp[1];
  t->nial[2] -= pStep / t->mfp[2];
  t->nial[3] -= pStep / t->mfp[3];
}

static int ApplyMeanEnergyLoss(const struct G4HepEmData* d, const struct G4HepEmParameters* p, oel* el) { /* :216-259 */
  const double pStep = el->pstep;
  otrack* t = &el->t;
  const int isEl = t->charge < 0.0;
  const double ekin = t->ekin;
  const double range = el->range;
  if (pStep >= range || ekin <= p->fMinLossTableEnergy) {
    t->edep = ekin;
    SetEKin(t, 0.0);
    return 1;
  }
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const int imc = t->imc;
  const double lekin = GetLogEKin(t);
  double eloss = pStep * GetRestDEDX(e, imc, ekin, lekin);
  const int ireg = d->fTheMatCutData->fMatCutData[imc].fG4RegionIndex;
  const double linLim = p->fParametersPerRegion[ireg].fLinELossLimit;
  if (eloss > ekin * linLim) {
    const double postRange = range - pStep;
    eloss = ekin - GetInvRange(e, imc, postRange);
  }
  eloss = dmax(eloss, 0.0);
  if (eloss >= ekin) {
    eloss = ekin;
    SetEKin(t, 0);
    t->edep = eloss;
    return 1;
  }
  SetEKin(t, ekin - eloss);
  t->edep = eloss;
  return 0;
}

static void SampleMSC(const struct G4HepEmData* d, const struct G4HepEmParameters* p, oel* el, orng* rng) { /* :261-322 */
  const double pStep = el->pstep;
  otrack* t = &el->t;
  const int isEl = t->charge < 0.0;
  const int imc = t->imc;
  const double preEkin = el->pre_ekin;
  const double range = el->range;
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const double kTLimitMinfix = 1.0E-8, kTauSmall = 1.0e-16;
  omsc* m = &el->msc;
  if (m->active && (pStep > dmax(kTLimitMinfix, kTauSmall * m->lambtr1))) {
    double postEkin = preEkin;
    double postLEkin = el->pre_lekin;
    if (pStep > range * 0.01) {  /* G4VERSION_NUM >= 1100 */
      postEkin = t->ekin;
      postLEkin = GetLogEKin(t);
    }
    const int imat = d->fTheMatCutData->fMatCutData[imc].fHepEmMatIndex;
    const double postTr1mfp = GetTransportMFP(e, imat, postEkin, postLEkin);
    UMSC_SampleScattering(d, m, pStep, preEkin, m->lambtr1, postEkin, postTr1mfp, imat, isEl, p->fIsMSCPositronCor,
                          p->fIsMSCDisplacement, rng);
    if (!m->no_scat) {
      RotateToReferenceFrame(m->dir, t->dir);
      if (m->is_displace) RotateToReferenceFrame(m->displ, t->dir);
      t->dir[0] = m->dir[0];
      t->dir[1] = m->dir[1];
      t->dir[2] = m->dir[2];
    }
  }
}

static int SampleLossFluctuations(const struct G4HepEmData* d, const struct G4HepEmParameters* p, oel* el, orng* rng) { /* :324-368 */
  const double pStep = el->pstep;
  otrack* t = &el->t;
  const int isEl = t->charge < 0.0;
  const int imc = t->imc;
  const double preEkin = el->pre_ekin;
  double finalEkin = t->ekin;
  double eloss = t->edep;
  (void)pStep;
  const int ireg = d->fTheMatCutData->fMatCutData[imc].fG4RegionIndex;
  const int isFluct = p->fParametersPerRegion[ireg].fIsELossFluctuation;
  const double kFluctParMinEnergy = 1.E-5;
  if (isFluct && eloss > kFluctParMinEnergy) {
    const struct G4HepEmMCCData* mc = &d->fTheMatCutData->fMatCutData[imc];
    const double elCut = mc->fSecElProdCutE;
    const double meanExE = d->fTheMaterialData->fMaterialData[mc->fHepEmMatIndex].fMeanExEnergy;
    const double tmax = isEl ? 0.5 * preEkin : preEkin;
    const double tcut = dmin(elCut, tmax);
    eloss = SampleEnergyLossFluctuation(tcut, meanExE, eloss, rng);
    eloss = dmax(eloss, 0.0);
    finalEkin = preEkin - eloss;
  }
  if (finalEkin <= p->fElectronTrackingCut) {
    eloss = preEkin;
    finalEkin = 0.0;
    SetEKin(t, finalEkin);
    t->edep = eloss;
    return 1;
  }
  SetEKin(t, finalEkin);
  t->edep = eloss;
  return 0;
}

static int PerformContinuous(const struct G4HepEmData* d, const struct G4HepEmParameters* p, oel* el, orng* rng) { /* :375-405 */
  el->pre_ekin = el->t.ekin;
  el->pre_lekin = GetLogEKin(&el->t);
  UpdatePStepLength(el);
  const double pStep = el->pstep;
  if (pStep <= 0.0) return 0;
  UpdateNumIALeft(el);
  if (ApplyMeanEnergyLoss(d, p, el)) return 1;
  SampleMSC(d, p, el, rng);
  return SampleLossFluctuations(d, p, el, rng);
}

static int CheckDelta(const struct G4HepEmData* d, otrack* t, double rand) { /* :408-423 */
  const int isEl = t->charge < 0.0;
  const struct G4HepEmElectronData* e = ElData(d, isEl);
  const int iDProc = t->winner;
  const int imc = t->imc;
  const int imat = d->fTheMatCutData->fMatCutData[imc].fHepEmMatIndex;
  const double ekin = t->ekin;
  const double lekin = GetLogEKin(t);
  const double mxsec = iDProc < 2 ? GetRestMacXSec(e, imc, ekin, lekin, iDProc == 0)
                     : (iDProc < 3 ? ComputeMacXsecAnnihilation(ekin, d->fTheMaterialData->fMaterialData[imat].fElectronDensity)
                                   : GetMacXSecNuclear(e, imat, ekin, lekin));
  return mxsec <= 0.0 || rand > mxsec * t->mfp[iDProc];
}

static void PerformDiscrete(const struct G4HepEmData* d, const struct G4HepEmParameters* p, otl* tl) { /* :425-459 */
  otrack* t = &tl->el.t;
  const int isEl = t->charge < 0.0;
  const int iDProc = t->winner;
  if (iDProc < 0 || t->on_boundary) return;
  t->nial[iDProc] = -1.0;
  if (CheckDelta(d, t, flat(tl->rng))) return;
  const double ekin = t->ekin;
  switch (iDProc) {
    case 0: IoniPerform(tl, d, isEl); break;
    case 1: BremPerform(tl, d, isEl, ekin < p->fElectronBremModelLim); break;
    case 2: AnnihilateInFlight(tl); break;
    default: break;
  }
}

static void ElectronHowFar(const struct G4HepEmData* d, const struct G4HepEmParameters* p, otl* tl) { /* :35-45,166-169 */
  otrack* t = &tl->el.t;
  for (int ip = 0; ip < 4; ++ip)
    if (t->nial[ip] <= 0.) t->nial[ip] = -HLog(flat(tl->rng));
  HowFarToDiscreteInteraction(d, p, &tl->el);
  HowFarToMSC(d, p, &tl->el, tl->rng);
}

static int ElectronPerform(const struct G4HepEmData* d, const struct G4HepEmParameters* p, otl* tl) { /* :461-483 */
  otrack* t = &tl->el.t;
  t->edep = 0;
  tl->el.pstep = t->gstep;
  const int isEl = t->charge < 0.0;
  if (t->gstep <= 0.) return 0;
  const int stopped = PerformContinuous(d, p, &tl->el, tl->rng);
  if (stopped) {
    if (!isEl) AnnihilateAtRest(tl);
    return 1;
  }
  PerformDiscrete(d, p, tl);
  return 0;
}

/* ---- one e-/e+ step as G4HepEmTrackingManager::TrackElectron takes it ---------------
   G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:408-635, with the caller's geometry a slab
   lo <= x <= hi across the x axis. HowFarToDiscreteInteraction once; then, while MSC
   limits the sub-step, the region sets fIsMultipleStepsInMSCTrans and no plane is
   reached: HowFarToMSC, the straight move to the next plane at most, UpdatePStepLength,
   UpdateNumIALeft, ApplyMeanEnergyLoss and SampleMSC with the MSC displacement applied
   under the safety rule (:527-566); the next sub-step is what is left of the discrete
   step (:569-594). Then the loss fluctuation on the summed mean loss from the step's
   pre-step energy (:599-606), and at-rest annihilation or PerformDiscrete (:618-650,
   with a lepto-nuclear process present: winner 3 resets its interaction length and
   draws the delta test). A step that crosses a plane out of the world ends there
   with no along-step physics of that sub-step (:463-472). */
#define SLAB_MAX_SUBSTEPS 65536  /* the reference's loop is unbounded */

typedef struct oslab {
  double x, lo, hi;   /* pre-step position and planes */
  int exit_lo, exit_hi;
  int side, left;     /* out: the step ended on hi (+1) or lo (-1); left the world */
  double geom_path;   /* out: summed geometrical sub-steps */
  double true_path;   /* out: totalTruePathLength */
  int nsub;           /* out: sub-steps */
} oslab;

static int ElectronSlabStep(const struct G4HepEmData* d, const struct G4HepEmParameters* p, otl* tl, oslab* g) {
  oel* el = &tl->el;
  otrack* t = &el->t;
  const int isEl = t->charge < 0.0;
  const int ireg = d->fTheMatCutData->fMatCutData[t->imc].fG4RegionIndex;
  int cont = p->fParametersPerRegion[ireg].fIsMultipleStepsInMSCTrans;
  double x = g->x;
  {
    const double saf = dmin(x - g->lo, g->hi - x);
    t->safety = t->on_boundary ? 0.0 : dmax(saf, 0.0);
  }
  const double preStepEkin = t->ekin;
  const double preStepLogEkin = GetLogEKin(t);
  for (int ip = 0; ip < 4; ++ip)
    if (t->nial[ip] <= 0.) t->nial[ip] = -HLog(flat(tl->rng));
  HowFarToDiscreteInteraction(d, p, el);
  const int iDProc = t->winner;
  double stepLeft = el->pstep;
  double totalTrue = 0.0, totalEloss = 0.0;
  int stopped = 0;
  el->pre_ekin = t->ekin;
  el->pre_lekin = GetLogEKin(t);
  g->side = 0; g->left = 0; g->geom_path = 0.0; g->nsub = 0;
  for (int sub = 1;; ++sub) {
    HowFarToMSC(d, p, el, tl->rng);
    if (t->winner != -2) cont = 0;
    const double physicalStep = t->gstep;
    const double ux = t->dir[0];
    double dist = DBL_MAX;
    int sd = 0;
    if (ux > 0.0) { dist = (g->hi - x) / ux; sd = 1; }
    else if (ux < 0.0) { dist = (g->lo - x) / ux; sd = -1; }
    const int limited = dist < physicalStep;
    const double finalStep = limited ? dist : physicalStep;
    if (limited) {
      x = sd > 0 ? g->hi : g->lo;
    } else {
      x = x + finalStep * ux;
      x = x < g->lo ? g->lo : (x > g->hi ? g->hi : x);
    }
    g->geom_path += finalStep;
    g->nsub = sub;
    if (limited) {
      cont = 0;
      g->side = sd;
      if ((sd > 0 && g->exit_hi) || (sd < 0 && g->exit_lo)) { g->left = 1; break; }
    }
    t->gstep = finalStep;
    t->on_boundary = limited;
    if (finalStep > 0) {
      do {
        UpdatePStepLength(el);
        const double pStep = el->pstep;
        totalTrue += pStep;
        if (pStep <= 0.0) break;
        UpdateNumIALeft(el);
        stopped = ApplyMeanEnergyLoss(d, p, el);
        totalEloss += t->edep;
        if (stopped) { cont = 0; break; }
        SampleMSC(d, p, el, tl->rng);
        if (!limited) {
          const double d0 = el->msc.displ[0], d1 = el->msc.displ[1], d2 = el->msc.displ[2];
          const double dl2 = d0 * d0 + d1 * d1 + d2 * d2;
          if (dl2 > 5.0e-8 * 5.0e-8) {
            const double dispR = sqrt(dl2);
            const double ps = 0.99 * dmax(dmin(x - g->lo, g->hi - x), 0.0);
            if (ps > 0.0 && dispR <= ps) x += d0;
            else if (dispR < ps) x += d0;
            else if (ps > 5.0e-8) x += d0 * (ps / dispR);
          }
        }
      } while (0);
      if (cont) {
        t->edep = 0.0;
        t->winner = iDProc;
        el->pre_ekin = t->ekin;
        el->pre_lekin = GetLogEKin(t);
        const double pStep = el->pstep;
        stepLeft -= pStep;
        el->pstep = stepLeft;
        t->gstep = stepLeft;
        el->range -= pStep;
      }
    }
    if (!cont || sub >= SLAB_MAX_SUBSTEPS) break;
  }
  g->x = x;
  g->true_path = totalTrue;
  t->edep = totalEloss;
  if (!stopped) {
    el->pre_ekin = preStepEkin;
    el->pre_lekin = preStepLogEkin;
    stopped = SampleLossFluctuations(d, p, el, tl->rng);
  }
  if (stopped) {
    if (!isEl) AnnihilateAtRest(tl);
  } else if (!g->left) {
    PerformDiscrete(d, p, tl);
  }
  return stopped;
}

typedef struct eslab_job {
  const struct G4HepEmData* d; const struct G4HepEmParameters* p; uint64_t seed;
  hepem_tracks* t; hepem_step_out* res; double* x; const double* lo; const double* hi; const uint8_t* exitf;
  int8_t* side; uint8_t* left; int32_t* nsub; int begin, end;
} eslab_job;

static void* eslab_worker(void* arg) {
  const eslab_job* J = (const eslab_job*)arg;
  hepem_tracks* T = J->t;
  hepem_step_out* res = J->res;
  const int n = T->n;
  for (int i = J->begin; i < J->end; ++i) {
    orng rng;
    rng.key[0] = (uint32_t)J->seed;
    rng.key[1] = (uint32_t)(J->seed >> 32);
    rng.id = (uint32_t)T->id[i];
    rng.hi = T->rng_stream ? T->rng_stream[i] : 0u;
    rng.ctr = T->rng_ctr[i];
    rng.has_gauss = T->rng_has_gauss[i];
    rng.gauss = T->rng_gauss[i];
    otl tl;
    memset(&tl, 0, sizeof(tl));
    tl.rng = &rng;
    otrack* t = &tl.el.t;
    TrackReset(t);
    MscReset(&tl.el.msc);
    t->charge = T->ptype[i] == HEPEM_ELECTRON ? -1.0 : 1.0;
    tl.el.msc.init_range = T->msc_init_range[i];
    tl.el.msc.dyn_range_factor = T->msc_dyn_range_factor[i];
    tl.el.msc.tlimit_min = T->msc_tlimit_min[i];
    tl.el.msc.first_step = T->msc_first_step[i];
    SetEKin(t, T->ekin[i]);
    for (int c = 0; c < 3; ++c) t->dir[c] = T->dir[c * n + i];
    t->imc = T->imc[i];
    t->on_boundary = T->on_boundary[i];
    for (int k = 0; k < 4; ++k) t->nial[k] = T->num_ia_left[k * n + i];
    oslab g;
    g.x = J->x[i]; g.lo = J->lo[i]; g.hi = J->hi[i];
    g.exit_lo = J->exitf[i] & 1; g.exit_hi = (J->exitf[i] >> 1) & 1;
    const int stopped = ElectronSlabStep(J->d, J->p, &tl, &g);
    T->ekin[i] = t->ekin;
    for (int c = 0; c < 3; ++c) T->dir[c * n + i] = t->dir[c];
    for (int k = 0; k < 4; ++k) T->num_ia_left[k * n + i] = t->nial[k];
    T->safety[i] = t->safety;
    T->on_boundary[i] = (uint8_t)t->on_boundary;
    T->msc_init_range[i] = tl.el.msc.init_range;
    T->msc_dyn_range_factor[i] = tl.el.msc.dyn_range_factor;
    T->msc_tlimit_min[i] = tl.el.msc.tlimit_min;
    T->msc_first_step[i] = (uint8_t)tl.el.msc.first_step;
    T->rng_ctr[i] = rng.ctr;
    T->rng_gauss[i] = rng.gauss;
    T->rng_has_gauss[i] = (uint8_t)rng.has_gauss;
    J->x[i] = g.x;
    J->side[i] = (int8_t)g.side;
    J->left[i] = (uint8_t)g.left;
    J->nsub[i] = g.nsub;
    res->true_step[i] = g.true_path;
    res->geom_step[i] = g.geom_path;
    res->edep[i] = t->edep;
    for (int c = 0; c < 3; ++c) res->displacement[c * n + i] = 0.0;
    res->winner[i] = t->winner;
    res->stopped[i] = (uint8_t)(t->ekin <= 0.0);  /* stopped or killed in the step */
    (void)stopped;
    int ns = 0;
    osec* list[4];
    for (int k = 0; k < tl.nsec_el; ++k) list[ns++] = &tl.sec_el[k];
    for (int k = 0; k < tl.nsec_gm; ++k) list[ns++] = &tl.sec_gm[k];
    res->nsec[i] = ns;
    for (int s = 0; s < 2; ++s) {
      res->sec_ptype[s * n + i] = s < ns ? (int8_t)list[s]->ptype : -1;
      res->sec_ekin[s * n + i] = s < ns ? list[s]->ekin : 0.0;
      for (int c = 0; c < 3; ++c) res->sec_dir[(c * 2 + s) * n + i] = s < ns ? list[s]->dir[c] : 0.0;
    }
  }
  return NULL;
}

int oracle_electron_slab_step(const struct G4HepEmData* d, const struct G4HepEmParameters* p, uint64_t seed,
                              hepem_tracks* t, hepem_step_out* res, double* x, const double* lo, const double* hi,
                              const uint8_t* exitf, int8_t* side, uint8_t* left, int32_t* nsub, int nthreads) {
  if (!d || !p || !t || !res || !x || !lo || !hi || !exitf || !side || !left || !nsub || t->n < 0) return -1;
  const int n = t->n;
  for (int i = 0; i < n; ++i)
    if (t->ptype[i] == HEPEM_GAMMA || t->imc[i] < 0 || t->imc[i] >= d->fTheMatCutData->fNumMatCutData || !(t->ekin[i] > 0.0))
      return -1;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (nthreads > n) nthreads = n > 0 ? n : 1;
  eslab_job jobs[256];
  pthread_t th[256];
  const int chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    eslab_job* J = &jobs[k];
    J->d = d; J->p = p; J->seed = seed; J->t = t; J->res = res; J->x = x; J->lo = lo; J->hi = hi; J->exitf = exitf;
    J->side = side; J->left = left; J->nsub = nsub;
    J->begin = k * chunk < n ? k * chunk : n;
    J->end = (k + 1) * chunk < n ? (k + 1) * chunk : n;
  }
  if (nthreads == 1) {
    eslab_worker(&jobs[0]);
    return 0;
  }
  for (int k = 0; k < nthreads; ++k) pthread_create(&th[k], NULL, eslab_worker, &jobs[k]);
  for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
  return 0;
}

/* ---- G4HepEmGammaManager.icc --------------------------------------------------- */
static double GetMacXSecPE(const struct G4HepEmData* d, int imat, double ekin) { /* :155-170 */
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[imat];
  int interval = 0;
  if (ekin >= md->fSandiaEnergies[0]) {
    for (int i = md->fNumOfSandiaIntervals - 1; i >= 0; i--) {
      if (ekin >= md->fSandiaEnergies[i]) { interval = i; break; }
    }
  }
  const double* c = &md->fSandiaCoefficients[4 * interval];
  const double inv = 1 / ekin;
  return inv * (c[0] + inv * (c[1] + inv * (c[2] + inv * c[3])));
}
static double GetTotalMacXSec(const struct G4HepEmData* d, ogm* g) { /* :108-153 */
  otrack* t = &g->t;
  const double ekin = t->ekin;
  const double lekin = GetLogEKin(t);
  const int imat = d->fTheMatCutData->fMatCutData[t->imc].fHepEmMatIndex;
  const struct G4HepEmGammaData* gd = d->fTheGammaData;
  if (ekin > gd->fEMax1) {
    const int istart = imat * gd->fDataPerMat + gd->fNumData0 + gd->fNumData1;
    return GetSplineLog4(gd->fEGridSize2, &gd->fMacXsecData[istart], ekin, lekin, gd->fLogEMin2, gd->fEILDelta2, 1);
  }
  if (ekin > gd->fEMax0) {
    const int istart = imat * gd->fDataPerMat + gd->fNumData0;
    double mx[2];
    GetLinearLog2All(gd->fEGridSize1, &gd->fMacXsecData[istart], ekin, lekin, gd->fLogEMin1, gd->fEILDelta1, mx);
    g->pe_mxsec = mx[1];
    return mx[0];
  }
  const int istart = imat * gd->fDataPerMat;
  const double comp = GetLinearLog(gd->fEGridSize0, &gd->fMacXsecData[istart], ekin, lekin, gd->fLogEMin0, gd->fEILDelta0);
  const double pe = dmax(0.0, GetMacXSecPE(d, imat, ekin));
  g->pe_mxsec = pe;
  return comp + pe;
}
static void SampleInteraction(const struct G4HepEmData* d, ogm* g, double urnd) { /* :187-219 */
  otrack* t = &g->t;
  const double ekin = t->ekin;
  const double tmfp = t->mfp[0];
  t->nial[0] = -1.0;
  const struct G4HepEmGammaData* gd = d->fTheGammaData;
  if (ekin > gd->fEMax1) {
    const double lekin = GetLogEKin(t);
    const int imat = d->fTheMatCutData->fMatCutData[t->imc].fHepEmMatIndex;
    const int istart = imat * gd->fDataPerMat + gd->fNumData0 + gd->fNumData1;
    double mxSec = 0.0, cProb = 0.0;
    int pid = 2;
    do {
      if (pid == 5) { ++pid; break; }  /* gamma-nuclear bin: value unused (see header) */
      mxSec = GetSplineLog4(gd->fEGridSize2, &gd->fMacXsecData[istart], ekin, lekin, gd->fLogEMin2, gd->fEILDelta2, pid);
      cProb += mxSec * tmfp;
      ++pid;
    } while (pid < 6 && urnd > cProb);
    t->winner = pid - 3;
    g->pe_mxsec = mxSec;
    return;
  }
  const double mxPE = g->pe_mxsec;
  t->winner = (urnd > tmfp * mxPE) ? 1 : 2;
}

/* ---- gamma interactions ------------------------------------------------------- */
static double ComptonSample(double primE, double* primDir, const double* orgDir, orng* rng) { /* Compton.icc:67-103 */
  const double kappa = primE * kInvElectronMassC2;
  const double eps0 = 1. / (1. + 2. * kappa);
  const double eps02 = eps0 * eps0;
  const double al1 = -HLog(eps0);
  const double al2 = al1 + 0.5 * (1. - eps02);
  double eps, eps2, gf, oneMinusCost, sint2, r[3];
  do {
    flatArray(rng, 3, r);
    if (al1 > al2 * r[0]) {
      eps = HExp(-al1 * r[1]);
      eps2 = eps * eps;
    } else {
      eps2 = eps02 + (1. - eps02) * r[1];
      eps = sqrt(eps2);
    }
    oneMinusCost = (1. - eps) / (eps * kappa);
    sint2 = oneMinusCost * (2. - oneMinusCost);
    gf = 1. - eps * sint2 / (1. + eps2);
  } while (gf < r[2]);
  const double cost = 1.0 - oneMinusCost;
  const double sint = sqrt(dmax(0., sint2));
  const double phi = k2Pi * flat(rng);
  primDir[0] = sint * cos(phi);
  primDir[1] = sint * sin(phi);
  primDir[2] = cost;
  RotateToReferenceFrame(primDir, orgDir);
  return primE * eps;
}
static void ComptonPerform(otl* tl) { /* Compton.icc:17-65 */
  otrack* t = &tl->gm.t;
  const double primE = t->ekin;
  const double thr = 0.0001;
  if (primE < thr) return;
  const double org[3] = {t->dir[0], t->dir[1], t->dir[2]};
  const double postE = ComptonSample(primE, t->dir, org, tl->rng);
  const double secE = primE - postE;
  double edep = 0.0;
  if (secE > thr) {
    osec* s = AddSecEl(tl, HEPEM_ELECTRON);
    s->dir[0] = primE * org[0] - postE * t->dir[0];
    s->dir[1] = primE * org[1] - postE * t->dir[1];
    s->dir[2] = primE * org[2] - postE * t->dir[2];
    const double norm = 1.0 / sqrt(s->dir[0] * s->dir[0] + s->dir[1] * s->dir[1] + s->dir[2] * s->dir[2]);
    s->dir[0] *= norm;
    s->dir[1] *= norm;
    s->dir[2] *= norm;
    s->ekin = secE;
  } else {
    edep += secE;
  }
  if (postE > thr) {
    SetEKin(t, postE);
  } else {
    edep += postE;
    SetEKin(t, 0.0);
  }
  t->edep = edep;
}

static int ConvSelectTargetAtom(const struct G4HepEmGammaData* gd, int imat, double ekin, double lekin, double urndn) { /* Conversion.icc:151-176 */
  const int start = gd->fElemSelectorConvStartIndexPerMat[imat];
  const double* data = &gd->fElemSelectorConvData[start];
  const int numData = gd->fElemSelectorConvEgridSize;
  const int numElem = (int)data[0];
  const double logE0 = gd->fElemSelectorConvLogMinEkin, invLD = gd->fElemSelectorConvEILDelta;
  const double* xdata = gd->fElemSelectorConvEgrid;
  const double xv = dmax(xdata[0], dmin(xdata[numData - 1], ekin));
  const int idx = (int)dmax(0.0, dmin((lekin - logE0) * invLD, numData - 2.0));
  const double x1 = xdata[idx], x2 = xdata[idx + 1];
  const double dl = x2 - x1;
  const double b = dmax(0., dmin(1., (xv - x1) / dl));
  const int i0 = idx * (numElem - 1) + 1;
  const int i1 = i0 + (numElem - 1);
  int ie = 0;
  while (ie < numElem - 1 && urndn > data[i0 + ie] + b * (data[i1 + ie] - data[i0 + ie])) ++ie;
  return ie;
}
static void ConvScreen12(double delta, double* f1, double* f2) { /* Conversion.icc:266-274 */
  if (delta > 1.4) {
    *f1 = 42.038 - 8.29 * HLog(delta + 0.958);
    *f2 = *f1;
  } else {
    *f1 = 42.184 - delta * (7.444 - 1.623 * delta);
    *f2 = 41.326 - delta * (5.848 - 0.902 * delta);
  }
}
static double ConvScreen1(double delta) { return (delta > 1.4) ? 42.038 - 8.29 * HLog(delta + 0.958) : 42.184 - delta * (7.444 - 1.623 * delta); }
static double ConvScreen2(double delta) { return (delta > 1.4) ? 42.038 - 8.29 * HLog(delta + 0.958) : 41.326 - delta * (5.848 - 0.902 * delta); }
static void ConvPhi12(double delta, double* phi1, double* phi2) { /* Conversion.icc:240-248 */
  if (delta > 1.4) {
    *phi1 = 21.0190 - 4.145 * HLog(delta + 0.958);
    *phi2 = *phi1;
  } else {
    *phi1 = 20.806 - delta * (3.190 - 0.5710 * delta);
    *phi2 = 20.234 - delta * (2.126 - 0.0903 * delta);
  }
}
static double ConvRateNoLPM(double normCond, double epsMin, double epsRange, double deltaFactor, double invF10, double invF20,
                            double fz, orng* rng) { /* Conversion.icc:179-200 */
  double r[3], greject = 0., eps = 0.;
  do {
    flatArray(rng, 3, r);
    if (normCond > r[0]) {
      eps = 0.5 - epsRange * pow(r[1], 1. / 3.);
      const double delta = deltaFactor / (eps * (1. - eps));
      greject = (ConvScreen1(delta) - fz) * invF10;
    } else {
      eps = epsMin + epsRange * r[1];
      const double delta = deltaFactor / (eps * (1. - eps));
      greject = (ConvScreen2(delta) - fz) * invF20;
    }
  } while (greject < r[2]);
  return eps;
}
static double ConvRateWithLPM(double normCond, double epsMin, double epsRange, double deltaFactor, double invF10, double invF20,
                              double fz, orng* rng, double eGamma, double lpmEnergy, const struct G4HepEmElemData* ed) { /* :203-237 */
  const double z23 = ed->fZet23, ilVarS1 = ed->fILVarS1, ilVarS1Cond = ed->fILVarS1Cond;
  double r[3], greject = 0., eps = 0.;
  do {
    flatArray(rng, 3, r);
    double fXi, fG, fPhi, phi1, phi2;
    if (normCond > r[0]) {
      eps = 0.5 - epsRange * pow(r[1], 1. / 3.);
      const double delta = deltaFactor / (eps * (1. - eps));
      ConvPhi12(delta, &phi1, &phi2);
      EvaluateLPMFunctions(&fXi, &fG, &fPhi, eGamma, eps * eGamma, lpmEnergy, z23, ilVarS1, ilVarS1Cond, 0.0, -1.0);
      greject = fXi * ((2. * fPhi + fG) * phi1 - fG * phi2 - fPhi * fz) * invF10;
    } else {
      eps = epsMin + epsRange * r[1];
      const double delta = deltaFactor / (eps * (1. - eps));
      ConvPhi12(delta, &phi1, &phi2);
      EvaluateLPMFunctions(&fXi, &fG, &fPhi, eGamma, eps * eGamma, lpmEnergy, z23, ilVarS1, ilVarS1Cond, 0.0, -1.0);
      greject = fXi * ((fPhi + 0.5 * fG) * phi1 + 0.5 * fG * phi2 - 0.5 * (fG + fPhi) * fz) * invF20;
    }
  } while (greject < r[2]);
  return eps;
}
static void ConvPerform(otl* tl, const struct G4HepEmData* d) { /* Conversion.icc:19-53,56-146 */
  otrack* t = &tl->gm.t;
  const double primE = t->ekin;
  if (primE < 2. * kElectronMassC2) return;
  const double lPrimE = GetLogEKin(t);
  const int imc = t->imc;
  orng* rng = tl->rng;
  /* SampleKinEnergies */
  const int imat = d->fTheMatCutData->fMatCutData[imc].fHepEmMatIndex;
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[imat];
  const int elemIndx = (md->fNumOfElement > 1) ? ConvSelectTargetAtom(d->fTheGammaData, imat, primE, lPrimE, flat(rng)) : 0;
  const int iZet = md->fElementVect[elemIndx];
  const double lpmEnr = kLPMconstant * md->fRadiationLength;
  const int zcl = iZet < d->fTheElementData->fMaxZet ? iZet : d->fTheElementData->fMaxZet;
  const struct G4HepEmElemData* ed = &d->fTheElementData->fElementData[zcl];
  const double eps0 = kElectronMassC2 / primE;
  double eps = 0.0;
  if (primE < 2.0) {
    eps = eps0 + (0.5 - eps0) * flat(rng);
  } else {
    const double deltaFactor = eps0 * 136. / ed->fZet13;
    const double deltaMin = 4. * deltaFactor;
    const double deltaMax = (primE < 50.0) ? ed->fDeltaMaxLow : ed->fDeltaMaxHigh;
    const double logZ13 = 0.333333 * ed->fLogZ;
    const double FZ = (primE < 50.0) ? 8. * logZ13 : 8. * (logZ13 + ed->fCoulomb);
    const double epsp = 0.5 - 0.5 * sqrt(1. - deltaMin / deltaMax);
    const double epsMin = dmax(eps0, epsp);
    const double epsRange = 0.5 - epsMin;
    double F10, F20;
    ConvScreen12(deltaMin, &F10, &F20);
    F10 -= FZ;
    F20 -= FZ;
    const double NormF1 = dmax(F10 * epsRange * epsRange, 0.);
    const double NormF2 = dmax(1.5 * F20, 0.);
    const double NormCond = NormF1 / (NormF1 + NormF2);
    eps = (primE < 100000.0) ? ConvRateNoLPM(NormCond, epsMin, epsRange, deltaFactor, 1. / F10, 1. / F20, FZ, rng)
                             : ConvRateWithLPM(NormCond, epsMin, epsRange, deltaFactor, 1. / F10, 1. / F20, FZ, rng, primE, lpmEnr, ed);
  }
  double eTot, pTot;
  if (flat(rng) > 0.5) {
    eTot = (1. - eps) * primE;
    pTot = eps * primE;
  } else {
    pTot = (1. - eps) * primE;
    eTot = eps * primE;
  }
  const double eKin = dmax(0., eTot - kElectronMassC2);
  const double pKin = dmax(0., pTot - kElectronMassC2);
  osec* se = AddSecEl(tl, HEPEM_ELECTRON);
  osec* sp = AddSecEl(tl, HEPEM_POSITRON);
  /* SampleDirections: Conversion.icc:123-146 */
  const double phi = k2Pi * flat(rng);
  const double cosPhi = cos(phi), sinPhi = sin(phi);
  const double costEl = SampleCostModifiedTsai(eKin, rng);
  const double sintEl = sqrt((1.0 - costEl) * (1.0 + costEl));
  se->dir[0] = sintEl * cosPhi;
  se->dir[1] = sintEl * sinPhi;
  se->dir[2] = costEl;
  RotateToReferenceFrame(se->dir, t->dir);
  const double costPos = SampleCostModifiedTsai(pKin, rng);
  const double sintPos = sqrt((1.0 - costPos) * (1.0 + costPos));
  sp->dir[0] = -sintPos * cosPhi;
  sp->dir[1] = -sintPos * sinPhi;
  sp->dir[2] = costPos;
  RotateToReferenceFrame(sp->dir, t->dir);
  se->ekin = eKin;
  sp->ekin = pKin;
  SetEKin(t, 0.0);
}

static double PESelectElementBindingEnergy(const struct G4HepEmData* d, int imc, double mxsec, double ekin, orng* rng) { /* Photoelectric.icc:39-78 */
  const int imat = d->fTheMatCutData->fMatCutData[imc].fHepEmMatIndex;
  const struct G4HepEmMatData* md = &d->fTheMaterialData->fMaterialData[imat];
  int ielem = 0;
  if (md->fNumOfElement > 1) {
    const double x = flat(rng) * mxsec;
    double sum = 0;
    const double invE = 1 / ekin;
    for (int i = 0; i < md->fNumOfElement; i++) {
      const struct G4HepEmElemData* ed = &d->fTheElementData->fElementData[md->fElementVect[i]];
      int interval = 0;
      if (ekin >= ed->fSandiaEnergies[0]) {
        for (int j = ed->fNumOfSandiaIntervals - 1; j >= 0; j--) {
          if (ekin >= ed->fSandiaEnergies[j]) { interval = j; break; }
        }
      }
      const double* c = &ed->fSandiaCoefficients[4 * interval];
      sum += md->fNumOfAtomsPerVolumeVect[i] * invE * (c[0] + invE * (c[1] + invE * (c[2] + invE * c[3])));
      if (x <= sum) { ielem = i; break; }
    }
  }
  return d->fTheElementData->fElementData[md->fElementVect[ielem]].fKShellBindingEnergy;
}
static void PEPerform(otl* tl, const struct G4HepEmData* d) { /* Photoelectric.icc:13-37,80-121 */
  otrack* t = &tl->gm.t;
  const double gE = t->ekin;
  const double be = PESelectElementBindingEnergy(d, t->imc, tl->gm.pe_mxsec, gE, tl->rng);
  const double thr = 0.000001;
  const double peE = gE - be;
  if (peE > thr) {
    osec* s = AddSecEl(tl, HEPEM_ELECTRON);
    const double tau = peE * kInvElectronMassC2;
    const double gamma = 1.0 + tau;
    const double beta = sqrt(tau * (tau + 2.0)) / gamma;
    const double ac = (1.0 - beta) / beta;
    const double a1 = 0.5 * beta * gamma * tau * (gamma - 2.0);
    const double a2 = ac + 2.0;
    const double gtmax = 2.0 * (a1 + 1.0 / ac);
    double tsam = 0.0, gtr = 0.0, r[2];
    do {
      flatArray(tl->rng, 2, r);
      tsam = 2.0 * ac * (2.0 * r[0] + a2 * sqrt(r[0])) / (a2 * a2 - 4.0 * r[0]);
      gtr = (2.0 - tsam) * (a1 + 1.0 / (ac + tsam));
    } while (r[1] * gtmax > gtr);
    const double costheta = 1.0 - tsam;
    const double sint = sqrt(tsam * (2.0 - tsam));
    const double phi = k2Pi * flat(tl->rng);
    s->dir[0] = sint * cos(phi);
    s->dir[1] = sint * sin(phi);
    s->dir[2] = costheta;
    RotateToReferenceFrame(s->dir, t->dir);
    s->ekin = peE;
    t->edep = be;
  } else {
    t->edep = gE;
  }
  SetEKin(t, 0.0);
}

static void GammaHowFar(const struct G4HepEmData* d, otl* tl) { /* GammaManager.icc:27-48 */
  otrack* t = &tl->gm.t;
  if (t->nial[0] <= 0.0) t->nial[0] = -HLog(flat(tl->rng));
  const double tot = GetTotalMacXSec(d, &tl->gm);
  const double mfp = (tot > 0.) ? 1. / tot : kALargeValue;
  t->mfp[0] = mfp;
  t->gstep = mfp * t->nial[0];
}
static void GammaPerform(const struct G4HepEmData* d, const struct G4HepEmParameters* p, otl* tl) { /* GammaManager.icc:54-105 */
  otrack* t = &tl->gm.t;
  t->nial[0] -= t->gstep / t->mfp[0];
  t->edep = 0.0;
  if (t->on_boundary) return;
  const int iDProc = t->winner;
  t->nial[iDProc] = -1.0;
  switch (iDProc) {
    case 0: ConvPerform(tl, d); break;
    case 1: ComptonPerform(tl); break;
    case 2: PEPerform(tl, d); break;
    default: break;
  }
  const double finalEkin = t->ekin;
  if (finalEkin > 0.0 && finalEkin <= p->fGammaTrackingCut) {
    SetEKin(t, 0.0);
    t->edep += finalEkin;
  }
}

/* ---- drivers ------------------------------------------------------------------------ */
int oracle_lookup_eloss_xsec(const struct G4HepEmData* d, int ptype, int n, const int32_t* imc, const double* ekin,
                             const double* lekin, double* out) {
  if (!d || n < 0 || (ptype != 0 && ptype != 1)) return -1;
  const struct G4HepEmElectronData* e = ElData(d, ptype == 0);
  for (int i = 0; i < n; ++i) {
    const double r = GetRestRange(e, imc[i], ekin[i], lekin[i]);
    out[i] = GetRestDEDX(e, imc[i], ekin[i], lekin[i]);
    out[n + i] = r;
    out[2 * n + i] = GetInvRange(e, imc[i], r);
    out[3 * n + i] = GetRestMacXSec(e, imc[i], ekin[i], lekin[i], 1);
    out[4 * n + i] = GetRestMacXSec(e, imc[i], ekin[i], lekin[i], 0);
  }
  return 0;
}

int oracle_probe_electron(const struct G4HepEmData* d, int ptype, int which, int n, const int32_t* imc, const double* ekin,
                          const double* lekin, double* out) {
  const struct G4HepEmElectronData* e = ElData(d, ptype == 0);
  for (int i = 0; i < n; ++i) {
    const int imat = d->fTheMatCutData->fMatCutData[imc[i]].fHepEmMatIndex;
    const double ne = d->fTheMaterialData->fMaterialData[imat].fElectronDensity;
    switch (which) {
      case 0: out[i] = GetRestMacXSecForStepping(e, imc[i], ekin[i], lekin[i], 1); break;
      case 1: out[i] = GetRestMacXSecForStepping(e, imc[i], ekin[i], lekin[i], 0); break;
      case 2: out[i] = GetTransportMFP(e, imat, ekin[i], lekin[i]); break;
      case 3: out[i] = GetMacXSecNuclear(e, imat, ekin[i], lekin[i]); break;
      case 4: out[i] = ComputeMacXsecAnnihilation(ekin[i], ne); break;
      case 5: out[i] = ComputeMacXsecAnnihilationForStepping(ekin[i], ne); break;
      default: return -1;
    }
  }
  return 0;
}

int oracle_probe_selector(const struct G4HepEmData* d, int ptype, int which, int n, const int32_t* imc, const double* ekin,
                          const double* lekin, const double* urnd, int32_t* out) {
  const struct G4HepEmElectronData* e = ElData(d, ptype == 0);
  for (int i = 0; i < n; ++i) {
    const int imat = d->fTheMatCutData->fMatCutData[imc[i]].fHepEmMatIndex;
    if (d->fTheMaterialData->fMaterialData[imat].fNumOfElement < 2) { out[i] = 0; continue; }
    switch (which) {
      case 0: out[i] = BremSelectTargetAtom(e, imc[i], ekin[i], lekin[i], urnd[i], 1); break;
      case 1: out[i] = BremSelectTargetAtom(e, imc[i], ekin[i], lekin[i], urnd[i], 0); break;
      case 2: out[i] = ConvSelectTargetAtom(d->fTheGammaData, imat, ekin[i], lekin[i], urnd[i]); break;
      default: return -1;
    }
  }
  return 0;
}

int oracle_probe_gamma(const struct G4HepEmData* d, int n, const int32_t* imc, const double* ekin, const double* lekin,
                       const double* urnd, double* out_xs, double* out_pe, int32_t* out_pid) {
  for (int i = 0; i < n; ++i) {
    ogm g;
    TrackReset(&g.t);
    g.pe_mxsec = 0.0;
    g.t.ekin = ekin[i];
    g.t.lekin = lekin[i];
    g.t.imc = imc[i];
    const double tot = GetTotalMacXSec(d, &g);
    g.t.mfp[0] = (tot > 0.) ? 1. / tot : kALargeValue;
    out_xs[i] = tot;
    SampleInteraction(d, &g, urnd[i]);
    out_pe[i] = g.pe_mxsec;
    out_pid[i] = g.t.winner;
  }
  return 0;
}

typedef struct step_job {
  const struct G4HepEmData* d;
  const struct G4HepEmParameters* p;
  uint64_t seed;
  const hepem_tracks* in;
  hepem_tracks* out;
  hepem_step_out* res;
  const double* geom_limit;
  int begin, end;
  /* optional transport of a gamma between HowFar and Perform in place of
     geom_limit (the slab shower's Woodcock tracking): returns 1 if it set the
     step (gstep, on_boundary and, for an interaction, couple, mfp and PE
     cross section), 0 to fall back to geom_limit */
  int (*gamma_transport)(const struct step_job* J, int i, otl* tl, void* user);
  void* user;
} step_job;

/* The caller's geometry between HowFar and Perform (G4HepEmTrackingManager.cc:
   262-300,1000-1092): the geometrical step is limited to geom_limit[i]; a
   limited step ends on a boundary, any other step does not. */
static void apply_geometry(const step_job* J, int i, otrack* t) {
  if (!J->geom_limit) return;
  if (J->geom_limit[i] < t->gstep) {
    t->gstep = J->geom_limit[i];
    t->on_boundary = 1;
  } else {
    t->on_boundary = 0;
  }
}

static void step_one(const step_job* J, int i) {
  const struct G4HepEmData* d = J->d;
  const struct G4HepEmParameters* p = J->p;
  const hepem_tracks* in = J->in;
  hepem_tracks* out = J->out;
  hepem_step_out* res = J->res;
  const int n = in->n;
  /* tracks with an invalid couple index or a non-positive/non-finite energy
     are passed through unchanged, flagged with winner -99 (library contract) */
  {
    const int imc = in->imc[i];
    const double e = in->ekin[i];
    if (!(imc >= 0 && imc < d->fTheMatCutData->fNumMatCutData && e > 0.0 && e < 1.0e300)) {
      res->true_step[i] = res->geom_step[i] = res->edep[i] = 0.0;
      for (int c = 0; c < 3; ++c) res->displacement[c * n + i] = 0.0;
      res->winner[i] = -99;
      res->stopped[i] = 1;
      res->nsec[i] = 0;
      for (int s = 0; s < 2; ++s) {
        res->sec_ptype[s * n + i] = -1;
        res->sec_ekin[s * n + i] = 0.0;
        for (int c = 0; c < 3; ++c) res->sec_dir[(c * 2 + s) * n + i] = 0.0;
      }
      if (out != in) {
        out->ptype[i] = in->ptype[i]; out->id[i] = in->id[i]; out->imc[i] = imc; out->ekin[i] = e;
        if (out->rng_stream) out->rng_stream[i] = in->rng_stream ? in->rng_stream[i] : 0u;
        for (int c = 0; c < 3; ++c) out->dir[c * n + i] = in->dir[c * n + i];
        out->safety[i] = in->safety[i]; out->on_boundary[i] = in->on_boundary[i];
        for (int k = 0; k < 4; ++k) out->num_ia_left[k * n + i] = in->num_ia_left[k * n + i];
        out->rng_ctr[i] = in->rng_ctr[i]; out->rng_gauss[i] = in->rng_gauss[i]; out->rng_has_gauss[i] = in->rng_has_gauss[i];
        if (in->ptype[i] != HEPEM_GAMMA) {
          out->msc_init_range[i] = in->msc_init_range[i]; out->msc_dyn_range_factor[i] = in->msc_dyn_range_factor[i];
          out->msc_tlimit_min[i] = in->msc_tlimit_min[i]; out->msc_first_step[i] = in->msc_first_step[i];
        }
      }
      return;
    }
  }
  orng rng;
  rng.key[0] = (uint32_t)J->seed;
  rng.key[1] = (uint32_t)(J->seed >> 32);
  rng.id = (uint32_t)in->id[i];
  rng.hi = in->rng_stream ? in->rng_stream[i] : 0u;
  rng.ctr = in->rng_ctr[i];
  rng.has_gauss = in->rng_has_gauss[i];
  rng.gauss = in->rng_gauss[i];
  otl tl;
  memset(&tl, 0, sizeof(tl));
  tl.rng = &rng;
  const int ptype = in->ptype[i];
  otrack* t;
  double displ[3] = {0., 0., 0.};
  double trueStep, geomStep;
  if (ptype == HEPEM_GAMMA) {
    t = &tl.gm.t;
    TrackReset(t);
    tl.gm.pe_mxsec = 0.0;
  } else {
    t = &tl.el.t;
    TrackReset(t);
    MscReset(&tl.el.msc);
    tl.el.range = 0.0;
    tl.el.pstep = 0.0;
    t->charge = ptype == HEPEM_ELECTRON ? -1.0 : 1.0;
    tl.el.msc.init_range = in->msc_init_range[i];
    tl.el.msc.dyn_range_factor = in->msc_dyn_range_factor[i];
    tl.el.msc.tlimit_min = in->msc_tlimit_min[i];
    tl.el.msc.first_step = in->msc_first_step[i];
  }
  SetEKin(t, in->ekin[i]);
  t->dir[0] = in->dir[i];
  t->dir[1] = in->dir[n + i];
  t->dir[2] = in->dir[2 * n + i];
  t->imc = in->imc[i];
  t->safety = in->safety[i];
  t->on_boundary = in->on_boundary[i];
  for (int k = 0; k < 4; ++k) t->nial[k] = in->num_ia_left[k * n + i];
  int stopped;
  if (ptype == HEPEM_GAMMA) {
    /* G4HepEmTrackingManager.cc:999-1000,1092-1108: HowFar, (no geometry),
       then UpdateNumIALeft on a boundary or SelectInteraction + Perform */
    GammaHowFar(d, &tl);
    if (!(J->gamma_transport && J->gamma_transport(J, i, &tl, J->user))) apply_geometry(J, i, t);
    if (t->on_boundary) {
      t->nial[0] -= t->gstep / t->mfp[0];
      t->edep = 0.0;
    } else {
      SampleInteraction(d, &tl.gm, flat(&rng));
      if (t->winner != 3) GammaPerform(d, p, &tl);
      else t->edep = 0.0;
    }
    trueStep = geomStep = t->gstep;
    stopped = t->ekin <= 0.0;
  } else {
    ElectronHowFar(d, p, &tl);
    apply_geometry(J, i, t);
    ElectronPerform(d, p, &tl);
    trueStep = tl.el.pstep;
    geomStep = t->gstep;
    displ[0] = tl.el.msc.displ[0];
    displ[1] = tl.el.msc.displ[1];
    displ[2] = tl.el.msc.displ[2];
    stopped = t->ekin <= 0.0;
    out->msc_init_range[i] = tl.el.msc.init_range;
    out->msc_dyn_range_factor[i] = tl.el.msc.dyn_range_factor;
    out->msc_tlimit_min[i] = tl.el.msc.tlimit_min;
    out->msc_first_step[i] = (uint8_t)tl.el.msc.first_step;
  }
  /* state out */
  out->ekin[i] = t->ekin;
  out->dir[i] = t->dir[0];
  out->dir[n + i] = t->dir[1];
  out->dir[2 * n + i] = t->dir[2];
  for (int k = 0; k < 4; ++k) out->num_ia_left[k * n + i] = t->nial[k];
  out->rng_ctr[i] = rng.ctr;
  out->rng_gauss[i] = rng.gauss;
  out->rng_has_gauss[i] = (uint8_t)rng.has_gauss;
  if (out != in) {
    out->ptype[i] = in->ptype[i];
    out->id[i] = in->id[i];
    if (out->rng_stream) out->rng_stream[i] = in->rng_stream ? in->rng_stream[i] : 0u;
    out->imc[i] = in->imc[i];
    out->safety[i] = in->safety[i];
  }
  out->on_boundary[i] = (uint8_t)t->on_boundary;
  /* results */
  res->true_step[i] = trueStep;
  res->geom_step[i] = geomStep;
  res->edep[i] = t->edep;
  res->displacement[i] = displ[0];
  res->displacement[n + i] = displ[1];
  res->displacement[2 * n + i] = displ[2];
  res->winner[i] = t->winner;
  res->stopped[i] = (uint8_t)stopped;
  /* secondaries: electron list first, then gammas (StackSecondaries order) */
  int ns = 0;
  osec* list[4];
  for (int k = 0; k < tl.nsec_el; ++k) list[ns++] = &tl.sec_el[k];
  for (int k = 0; k < tl.nsec_gm; ++k) list[ns++] = &tl.sec_gm[k];
  res->nsec[i] = ns;
  for (int s = 0; s < 2; ++s) {
    if (s < ns) {
      res->sec_ptype[s * n + i] = (int8_t)list[s]->ptype;
      res->sec_ekin[s * n + i] = list[s]->ekin;
      for (int c = 0; c < 3; ++c) res->sec_dir[(c * 2 + s) * n + i] = list[s]->dir[c];
    } else {
      res->sec_ptype[s * n + i] = -1;
      res->sec_ekin[s * n + i] = 0.0;
      for (int c = 0; c < 3; ++c) res->sec_dir[(c * 2 + s) * n + i] = 0.0;
    }
  }
}

static void* step_worker(void* arg) {
  const step_job* J = (const step_job*)arg;
  for (int i = J->begin; i < J->end; ++i) step_one(J, i);
  return NULL;
}

int oracle_step(const struct G4HepEmData* d, const struct G4HepEmParameters* p, uint64_t seed, const hepem_tracks* in,
                hepem_tracks* out, hepem_step_out* res, int nthreads) {
  return oracle_step_geom(d, p, seed, in, out, res, NULL, nthreads);
}

int oracle_step_geom(const struct G4HepEmData* d, const struct G4HepEmParameters* p, uint64_t seed, const hepem_tracks* in,
                     hepem_tracks* out, hepem_step_out* res, const double* geom_limit, int nthreads) {
  if (!d || !p || !in || !out || !res || in->n < 0) return -1;
  const int n = in->n;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (nthreads > n) nthreads = n > 0 ? n : 1;
  step_job jobs[256];
  pthread_t th[256];
  const int chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    jobs[k].d = d; jobs[k].p = p; jobs[k].seed = seed;
    jobs[k].in = in; jobs[k].out = out; jobs[k].res = res; jobs[k].geom_limit = geom_limit;
    jobs[k].gamma_transport = NULL; jobs[k].user = NULL;
    jobs[k].begin = k * chunk < n ? k * chunk : n;
    jobs[k].end = (k + 1) * chunk < n ? (k + 1) * chunk : n;
  }
  if (nthreads == 1) {
    step_worker(&jobs[0]);
    return 0;
  }
  for (int k = 0; k < nthreads; ++k) pthread_create(&th[k], NULL, step_worker, &jobs[k]);
  for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
  return 0;
}

/* ---- slab shower (CPU harness of hepem_slab_run) ------------------------------------
   The same tracking loop as the device driver (g4hepem_amd_shower.h): safety, HowFar,
   geometry cut at the next plane, Perform, MSC displacement with the safety rule of
   G4HepEmTrackingManager.cc:527-566, slab crossing, fixed-point scoring, and the same
   64-bit stream ids; tracks of an event are followed depth-first from a stack. */

#define SH_FIX 1.0e9
#define SH_GEOM_MIN 5.0e-8

static uint64_t sh_mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t sh_primary_stream(int event) { return sh_mix64(0x5EED000000000000ull ^ (uint64_t)event); }
static uint64_t sh_child_stream(uint64_t parent, int nstep, int slot) {
  return sh_mix64(parent ^ sh_mix64(((uint64_t)(uint32_t)nstep << 2) | (uint64_t)(slot + 1)));
}

typedef struct sh_track {
  int8_t ptype; int32_t id; uint32_t hi; int32_t imc; double ekin, dir[3], safety; uint8_t onb;
  double nial[4], mir, mdrf, mtm; uint8_t mfs; uint32_t ctr; double gauss; uint8_t hg;
  double x; int layer, nstep;
} sh_track;

/* Woodcock tracking of a gamma through the whole stack of slabs
   (G4HepEm/G4HepEm/src/G4HepEmWoodcockHelper.cc:145-300, driven at
   G4HepEmTrackingManager.cc:955-1062): steps of the densest absorber's mean
   free path until a point is accepted as an interaction (always in the densest
   material, else with probability Sigma/Sigma_max) or the exit of the stack,
   less 1e-3 mm, is reached. The event loop sets the pre-step point; the hook
   reports where the step ended. */
typedef struct sh_wdt {
  const double* plane; int nbins, nabs; const int* imc; int mat[HEPEM_SLAB_MAX_ABS];
  int wdt_imc, wdt_mat; double wdt_min, xmax;
  double x; int layer;                          /* pre-step point */
  int done, boundary, layer_out; double x_out;  /* result of the last step */
} sh_wdt;

static int sh_locate(const double* plane, int nbins, double x) {
  int lo = 0, hi = nbins;  /* plane[lo] <= x < plane[hi] */
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (x < plane[mid]) hi = mid; else lo = mid;
  }
  return lo;
}

static int sh_woodcock(const step_job* J, int i, otl* tl, void* user) {
  (void)i;
  sh_wdt* W = (sh_wdt*)user;
  otrack* t = &tl->gm.t;
  W->done = 0;
  if (!(W->wdt_min > 0.0) || !(t->ekin > W->wdt_min)) return 0;
  const double ux = t->dir[0];
  const double distOut = ux > 0.0 ? (W->xmax - W->x) / ux : (ux < 0.0 ? (0.0 - W->x) / ux : 1.0e300);
  double dist = distOut - 1.0E-3 > 0.0 ? distOut - 1.0E-3 : 0.0;
  if (!(dist >= 1.0E-6)) return 0;
  t->imc = W->wdt_imc;
  tl->gm.pe_mxsec = 0.0;
  const double wmx = GetTotalMacXSec(J->d, &tl->gm);
  const double wmfp = wmx > 0.0 ? 1.0 / wmx : DBL_MAX;
  const double wpe = tl->gm.pe_mxsec;
  double L = 0.0, mx = 0.0, mpe = 0.0, mfp = wmfp, pint = wpe;
  int prev = -1, lp = W->layer, imcp = W->wdt_imc, boundary = 0;
  for (int it = 0;; ++it) {
    if (it >= (1 << 20)) { L += dist; boundary = 1; break; }
    const double pstep = wmfp < DBL_MAX ? -HLog(flat(tl->rng)) * wmfp : DBL_MAX;
    if (dist < pstep) { L += dist; boundary = 1; break; }
    L += pstep;
    dist -= pstep;
    const double xp = W->x + L * ux;
    lp = sh_locate(W->plane, W->nbins, xp);
    const int a = lp % W->nabs;
    if (W->mat[a] != W->wdt_mat) {
      if (W->imc[a] != prev) {
        prev = W->imc[a];
        t->imc = prev;
        tl->gm.pe_mxsec = 0.0;
        mx = GetTotalMacXSec(J->d, &tl->gm);
        mpe = tl->gm.pe_mxsec;
      }
      if (mx * wmfp > flat(tl->rng)) { imcp = W->imc[a]; mfp = mx > 0.0 ? 1.0 / mx : DBL_MAX; pint = mpe; break; }
    } else {
      /* in the densest material: an interaction for sure, with its couple */
      imcp = W->wdt_imc; mfp = wmfp; pint = wpe;
      break;
    }
  }
  t->nial[0] = -1.0;
  t->gstep = 0.0;
  W->done = 1;
  W->boundary = boundary;
  if (boundary) {
    t->on_boundary = 1;
  } else {
    t->on_boundary = 0;
    t->imc = imcp;
    t->mfp[0] = mfp;
    tl->gm.pe_mxsec = pint;
    W->layer_out = lp;
    W->x_out = W->x + L * ux;
  }
  return 1;
}

typedef struct sh_job {
  const struct G4HepEmData* d; const struct G4HepEmParameters* p; uint64_t seed;
  const hepem_slab_setup* su; const double* plane; int nbins;
  sh_wdt wdt;          /* Woodcock set-up (per-step fields are per event) */
  int e_begin, e_end;
  int64_t* dep;        /* [nevents][nbins] fixed point */
  int64_t* leak;       /* [nevents] */
  int64_t steps, tracks;
  int status;
} sh_job;

static void sh_event(sh_job* J, int ev) {
  const int nbins = J->nbins, nabs = J->su->nabs;
  const double* P = J->plane;
  int cap = 4096, top = 0;
  sh_track* stk = (sh_track*)malloc((size_t)cap * sizeof(sh_track));
  if (!stk) { J->status = -4; return; }
  {
    sh_track t;
    memset(&t, 0, sizeof(t));
    const uint64_t sid = sh_primary_stream(ev);
    t.ptype = (int8_t)J->su->primary; t.id = (int32_t)(uint32_t)sid; t.hi = (uint32_t)(sid >> 32);
    t.imc = J->su->imc[0]; t.ekin = J->su->ekin0; t.dir[0] = 1.0; t.onb = 1;
    for (int k = 0; k < 4; ++k) t.nial[k] = -1.0;
    t.mir = 1.0e+21; t.mdrf = 0.04; t.mtm = 1.0E-7; t.mfs = 1;
    stk[top++] = t;
    J->tracks++;
  }
  /* one-track SoA views for the gamma steps (step_one) */
  int8_t a_pt; int32_t a_id, a_imc; uint32_t a_hi, a_ctr; double a_e, a_dir[3], a_saf, a_nial[4], a_mir, a_mdrf, a_mtm, a_g;
  uint8_t a_onb, a_mfs, a_hg;
  double r_ts, r_gs, r_ed, r_disp[3], r_sek[2], r_sdir[6]; int32_t r_w, r_ns; uint8_t r_st; int8_t r_spt[2];
  hepem_tracks T = {1, &a_pt, &a_id, &a_imc, &a_e, a_dir, &a_saf, &a_onb, a_nial, &a_mir, &a_mdrf, &a_mtm, &a_mfs,
                    &a_ctr, &a_g, &a_hg, &a_hi};
  hepem_step_out R = {&r_ts, &r_gs, &r_ed, r_disp, &r_w, &r_st, &r_ns, r_spt, r_sek, r_sdir};
  double glim;
  sh_wdt W = J->wdt;
  step_job SJ;
  SJ.d = J->d; SJ.p = J->p; SJ.seed = J->seed; SJ.in = &T; SJ.out = &T; SJ.res = &R; SJ.geom_limit = &glim;
  SJ.begin = 0; SJ.end = 1;
  SJ.gamma_transport = W.wdt_min > 0.0 ? sh_woodcock : NULL; SJ.user = &W;
  while (top > 0) {
    sh_track t = stk[--top];
    for (;;) {
      int l = t.layer;
      const double lo = P[l], hi = P[l + 1];
      double x;            /* post-step position */
      int onb, sd;         /* post-step boundary flag; +1/-1: upper/lower plane */
      int left = 0;        /* an e-/e+ that left the world (the stack) in the step */
      double ed, ekin;     /* the step's deposit; post-step energy */
      int ns = 0;
      int8_t spt[2] = {-1, -1};
      double sek[2] = {0., 0.}, sdir[6] = {0., 0., 0., 0., 0., 0.};
      int l0 = l;          /* the pre-step slab: scoring and the couple's cuts */
      if (t.ptype != 2) {
        /* e-/e+: the whole TrackElectron step, sub-steps and MSC displacements included */
        orng rng;
        rng.key[0] = (uint32_t)J->seed; rng.key[1] = (uint32_t)(J->seed >> 32);
        rng.id = (uint32_t)t.id; rng.hi = t.hi; rng.ctr = t.ctr; rng.has_gauss = t.hg; rng.gauss = t.gauss;
        otl tl;
        memset(&tl, 0, sizeof(tl));
        tl.rng = &rng;
        otrack* e = &tl.el.t;
        TrackReset(e);
        MscReset(&tl.el.msc);
        e->charge = t.ptype == HEPEM_ELECTRON ? -1.0 : 1.0;
        tl.el.msc.init_range = t.mir; tl.el.msc.dyn_range_factor = t.mdrf; tl.el.msc.tlimit_min = t.mtm;
        tl.el.msc.first_step = t.mfs;
        SetEKin(e, t.ekin);
        for (int c = 0; c < 3; ++c) e->dir[c] = t.dir[c];
        e->imc = t.imc;
        e->on_boundary = t.onb;
        for (int k = 0; k < 4; ++k) e->nial[k] = t.nial[k];
        oslab g;
        g.x = t.x; g.lo = lo; g.hi = hi; g.exit_lo = l == 0; g.exit_hi = l == nbins - 1;
        ElectronSlabStep(J->d, J->p, &tl, &g);
        x = g.x; sd = g.side; onb = sd != 0; left = g.left;
        ed = e->edep; ekin = e->ekin;
        for (int k = 0; k < tl.nsec_el + tl.nsec_gm && ns < 2; ++k, ++ns) {
          const osec* sc = k < tl.nsec_el ? &tl.sec_el[k] : &tl.sec_gm[k - tl.nsec_el];
          spt[ns] = (int8_t)sc->ptype; sek[ns] = sc->ekin;
          for (int c = 0; c < 3; ++c) sdir[c * 2 + ns] = sc->dir[c];
        }
        for (int c = 0; c < 3; ++c) t.dir[c] = e->dir[c];
        for (int k = 0; k < 4; ++k) t.nial[k] = e->nial[k];
        t.mir = tl.el.msc.init_range; t.mdrf = tl.el.msc.dyn_range_factor; t.mtm = tl.el.msc.tlimit_min;
        t.mfs = (uint8_t)tl.el.msc.first_step;
        t.ctr = rng.ctr; t.gauss = rng.gauss; t.hg = (uint8_t)rng.has_gauss;
      } else {
        /* gamma: HowFar, the cut at the next plane (or a Woodcock step), Perform */
        double saf = t.x - lo < hi - t.x ? t.x - lo : hi - t.x;
        t.safety = saf > 0.0 ? saf : 0.0;
        a_pt = t.ptype; a_id = t.id; a_hi = t.hi; a_imc = t.imc; a_e = t.ekin;
        a_dir[0] = t.dir[0]; a_dir[1] = t.dir[1]; a_dir[2] = t.dir[2]; a_saf = t.safety; a_onb = t.onb;
        for (int k = 0; k < 4; ++k) a_nial[k] = t.nial[k];
        a_mir = t.mir; a_mdrf = t.mdrf; a_mtm = t.mtm; a_mfs = t.mfs; a_ctr = t.ctr; a_g = t.gauss; a_hg = t.hg;
        const double ux = t.dir[0];
        sd = 0;
        glim = 1.0e300;
        if (ux > 0.0) { glim = (hi - t.x) / ux; sd = 1; }
        else if (ux < 0.0) { glim = (lo - t.x) / ux; sd = -1; }
        W.x = t.x; W.layer = l; W.done = 0;
        step_one(&SJ, 0);
        onb = a_onb;
        if (W.done) {
          /* Woodcock step: the pre-step point moves to where the step ended, an
             interaction point or the stack's last slab before its exit */
          if (W.boundary) {
            sd = ux > 0.0 ? 1 : (ux < 0.0 ? -1 : 1);
            l = sd > 0 ? nbins - 1 : 0;
          } else {
            l = W.layer_out;
          }
        }
        l0 = l;
        if (W.done && !W.boundary) x = W.x_out;
        else if (onb) x = sd > 0 ? P[l + 1] : P[l];
        else {
          x = t.x + r_gs * ux;
          x = x < lo ? lo : (x > hi ? hi : x);
        }
        if (!onb) sd = 0;
        ed = r_ed; ekin = a_e;
        ns = r_ns;
        for (int k = 0; k < ns; ++k) {
          spt[k] = r_spt[k]; sek[k] = r_sek[k];
          for (int c = 0; c < 3; ++c) sdir[c * 2 + k] = r_sdir[c * 2 + k];
        }
        t.dir[0] = a_dir[0]; t.dir[1] = a_dir[1]; t.dir[2] = a_dir[2];
        for (int k = 0; k < 4; ++k) t.nial[k] = a_nial[k];
        t.ctr = a_ctr; t.gauss = a_g; t.hg = a_hg;
      }
      J->steps++;
      int alive = ekin > 0.0;
      if (left || (alive && onb && (l + sd < 0 || l + sd >= nbins))) {
        /* leaves the stack: its kinetic energy, the annihilation energy of an e+
           in flight and anything its last step created are lost */
        double e = ekin + (t.ptype == 1 && ekin > 0.0 ? 2.0 * 0.51099895000 : 0.0);
        for (int k = 0; k < ns; ++k) e += sek[k];
        J->leak[ev] += (int64_t)llrint(e * SH_FIX);
        alive = 0; left = 1; ns = 0;
      } else if (alive && onb) {
        l += sd;
      }
      const uint64_t sid = ((uint64_t)t.hi << 32) | (uint32_t)t.id;
      const int ns1 = t.nstep + 1;
      /* StackSecondaries (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:1274-1313):
         where the couple's region applies cuts, a secondary below its production
         threshold is deposited at the step instead of stacked */
      int stacked[2] = {1, 1};
      {
        const struct G4HepEmMCCData* mcc = &J->d->fTheMatCutData->fMatCutData[J->su->imc[l0 % nabs]];
        const int apply = J->p->fParametersPerRegion[mcc->fG4RegionIndex].fIsApplyCuts;
        for (int s = 0; s < ns; ++s) {
          int cut = 0;
          if (apply) {
            if (spt[s] == 0) cut = sek[s] < mcc->fSecElProdCutE;
            else if (spt[s] == 1) cut = 0.51099895000 < mcc->fSecGamProdCutE && sek[s] < mcc->fSecPosProdCutE;
            else cut = sek[s] < mcc->fSecGamProdCutE;
          }
          if (cut) {
            ed += spt[s] == 1 ? sek[s] + 2.0 * 0.51099895000 : sek[s];
            stacked[s] = 0;
          }
        }
      }
      /* scoring */
      if (ed > 0.0) J->dep[(size_t)ev * nbins + l0] += (int64_t)llrint(ed * SH_FIX);
      /* secondaries: pushed in reverse so that slot 0 is followed first */
      if (top + ns + 1 > cap) {
        cap *= 2;
        sh_track* nstk = (sh_track*)realloc(stk, (size_t)cap * sizeof(sh_track));
        if (!nstk) { free(stk); J->status = -4; return; }
        stk = nstk;
      }
      for (int s = ns - 1; s >= 0; --s) {
        if (!stacked[s]) continue;
        sh_track c;
        memset(&c, 0, sizeof(c));
        const uint64_t cid = sh_child_stream(sid, ns1, s);
        c.ptype = spt[s]; c.id = (int32_t)(uint32_t)cid; c.hi = (uint32_t)(cid >> 32);
        c.imc = J->su->imc[l % nabs]; c.ekin = sek[s];
        for (int k = 0; k < 3; ++k) c.dir[k] = sdir[k * 2 + s];
        for (int k = 0; k < 4; ++k) c.nial[k] = -1.0;
        c.mir = 1.0e+21; c.mdrf = 0.04; c.mtm = 1.0E-7; c.mfs = 1;
        c.x = x; c.layer = l; c.nstep = 0;
        stk[top++] = c;
        J->tracks++;
      }
      if (!alive) break;
      /* the track goes on */
      t.ekin = ekin; t.onb = (uint8_t)onb;
      t.x = x; t.layer = l; t.imc = J->su->imc[l % nabs]; t.nstep = ns1;
    }
  }
  free(stk);
}

static void* sh_worker(void* arg) {
  sh_job* J = (sh_job*)arg;
  for (int e = J->e_begin; e < J->e_end && J->status == 0; ++e) sh_event(J, e);
  return NULL;
}

int oracle_slab_shower(const struct G4HepEmData* d, const struct G4HepEmParameters* p, const hepem_slab_setup* su,
                       int nevents, uint64_t seed, int64_t* dep, int64_t* leak, int64_t* steps_tracks, int nthreads) {
  if (!d || !p || !su || !dep || !leak || nevents < 0 || su->nabs < 1 || su->nabs > HEPEM_SLAB_MAX_ABS || su->nlayers < 1)
    return -1;
  const int nbins = su->nlayers * su->nabs;
  double* plane = (double*)malloc((size_t)(nbins + 1) * sizeof(double));
  if (!plane) return -4;
  plane[0] = 0.0;
  for (int k = 0; k < nbins; ++k) plane[k + 1] = plane[k] + su->thickness[k % su->nabs];
  memset(dep, 0, (size_t)nevents * nbins * sizeof(int64_t));
  memset(leak, 0, (size_t)nevents * sizeof(int64_t));
  /* Woodcock set-up: the densest material of the stack, first absorber on a tie
     (G4HepEmWoodcockHelper::FindWDTMaterial, G4HepEmWoodcockHelper.cc:318-338) */
  sh_wdt W;
  memset(&W, 0, sizeof(W));
  W.plane = plane; W.nbins = nbins; W.nabs = su->nabs; W.imc = su->imc;
  W.wdt_min = su->woodcock_min_ekin; W.xmax = plane[nbins]; W.wdt_imc = su->imc[0]; W.wdt_mat = -1;
  {
    double maxDensity = -1.0;
    for (int a = 0; a < su->nabs; ++a) {
      const int imat = d->fTheMatCutData->fMatCutData[su->imc[a]].fHepEmMatIndex;
      const double density = d->fTheMaterialData->fMaterialData[imat].fDensity;
      W.mat[a] = imat;
      if (density > maxDensity) { maxDensity = density; W.wdt_mat = imat; W.wdt_imc = su->imc[a]; }
    }
  }
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (nthreads > nevents) nthreads = nevents > 0 ? nevents : 1;
  sh_job jobs[256];
  pthread_t th[256];
  for (int k = 0; k < nthreads; ++k) {
    sh_job* J = &jobs[k];
    J->d = d; J->p = p; J->seed = seed; J->su = su; J->plane = plane; J->nbins = nbins; J->wdt = W;
    J->e_begin = (int)((int64_t)nevents * k / nthreads);
    J->e_end = (int)((int64_t)nevents * (k + 1) / nthreads);
    J->dep = dep; J->leak = leak; J->steps = 0; J->tracks = 0; J->status = 0;
  }
  for (int k = 0; k < nthreads; ++k) pthread_create(&th[k], NULL, sh_worker, &jobs[k]);
  int status = 0;
  int64_t steps = 0, tracks = 0;
  for (int k = 0; k < nthreads; ++k) {
    pthread_join(th[k], NULL);
    if (jobs[k].status) status = jobs[k].status;
    steps += jobs[k].steps;
    tracks += jobs[k].tracks;
  }
  free(plane);
  if (steps_tracks) { steps_tracks[0] = steps; steps_tracks[1] = tracks; }
  return status;
}
