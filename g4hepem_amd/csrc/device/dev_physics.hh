// dev_physics.hh -- device implementation of the G4HepEm run-time physics:
// log-grid spline lookups, element selectors, interaction samplers, Urban MSC,
// energy-loss fluctuation and the e-/e+ and gamma managers' HowFar/Perform.
//
// Each function names the reference function it reproduces (paths relative to
// G4HepEm/G4HepEmRun/include of the reference tree). Templated on the
// arithmetic/table type R: R = double is the FP64 form (same operations in the
// same order as the reference host code), R = float the FP32-table form, where
// the marked FP64 islands stay in double.
#pragma once

#include "dev_common.hh"

namespace hepem {

// ---- G4HepEmRunUtils.icc -------------------------------------------------------------
template <typename R> HD void Rotate(R* dir, const R* ref) {  // RotateToReferenceFrame :32-46
  R up = ref[0] * ref[0] + ref[1] * ref[1];
  if (up > (R)0.) {
    up = dsqrt(up);
    const R px = dir[0], py = dir[1], pz = dir[2];
    dir[0] = (ref[0] * ref[2] * px - ref[1] * py) / up + ref[0] * pz;
    dir[1] = (ref[1] * ref[2] * px + ref[0] * py) / up + ref[1] * pz;
    dir[2] = -up * px + ref[2] * pz;
  } else if (ref[2] < (R)0.) {
    dir[0] = -dir[0];
    dir[2] = -dir[2];
  }
}
// Interpolation: abscissas x, x1, x2 are double in both forms (see DevEl),
// ordinates and second derivatives have the form's precision R.
// Fraction b = (x - x1) / (x2 - x1) clamped to [0, 1]: both differences in
// double (they cancel near thresholds); the quotient in the form's precision
// (the FP64 form evaluates the reference's expression as written).
HD double Frac(double dx, double dl, double) { return tmax(0., tmin(1., dx / dl)); }
HD float Frac(double dx, double dl, float) { return tmax(0.f, tmin(1.f, (float)dx / (float)dl)); }
template <typename R> HD R Spline(double x1, double x2, R y1, R y2, R s1, R s2, double x) {  // GetSpline :49-61
  const double dl = x2 - x1;
  const R b = Frac(x - x1, dl, R());
  const R os = (R)0.166666666667;
  const R c0 = ((R)2.0 - b) * s1;
  const R c1 = ((R)1.0 + b) * s2;
  return y1 + b * (y2 - y1) + (b * (b - (R)1.0)) * (c0 + c1) * ((R)dl * (R)dl * os);
}
template <typename R> HD R Linear(double x1, double x2, R y1, R y2, double x) {  // GetLinear :63-71
  const double dl = x2 - x1;
  const R b = Frac(x - x1, dl, R());
  return y1 + b * (y2 - y1);
}
template <typename R> HD int LogBin(R logx, R logxmin, R invLD, int ndata) {
  return (int)tmax((R)0., tmin((logx - logxmin) * invLD, (R)(ndata - 2)));
}
// x grid separate, [y, y''] interleaved (GetSplineLog :84-91); xlo, xhi are
// the grid's first and last values (xd[0], xd[nd-1]), held by the caller
template <typename R>
HD R SplineLog2(int nd, const double* xd, const R* yd, double x, R logx, R lmin, R ild, double xlo, double xhi) {
  const double xv = tmax(xlo, tmin(xhi, x));
  const int idx = LogBin(logx, lmin, ild, nd);
  const int i2 = 2 * idx;
  return Spline(xd[idx], xd[idx + 1], yd[i2], yd[i2 + 2], yd[i2 + 1], yd[i2 + 3], xv);
}
// [x, y, y''] interleaved (GetSplineLog :95-102); x read from the double copy
// dx; xlo, xhi are the first and last abscissas (dx[0], dx[3(nd-1)])
template <typename R>
HD R SplineLog3(int nd, const double* dx, const R* d, double x, R logx, R lmin, R ild, double xlo, double xhi) {
  const double xv = tmax(xlo, tmin(xhi, x));
  const int idx = LogBin(logx, lmin, ild, nd);
  const int i3 = 3 * idx;
  return Spline(dx[i3], dx[i3 + 3], d[i3 + 1], d[i3 + 4], d[i3 + 2], d[i3 + 5], xv);
}
template <typename R> HD R LinearLog(int nd, const double* dx, const R* d, double x, R logx, R lmin, R ild) {  // :125-133
  const int idx = LogBin(logx, lmin, ild, nd);
  const int i0 = 2 * idx, i1 = i0 + 2;
  return Linear(dx[i0], dx[i1], d[i0 + 1], d[i1 + 1], x);
}
template <typename R>
HD void LinearLog2All(int nd, const double* dx, const R* d, double x, R logx, R lmin, R ild, R res[2]) {  // :146-161
  const int idx = LogBin(logx, lmin, ild, nd);
  const int i0 = 3 * idx, i1 = i0 + 3;
  const double dl = dx[i1] - dx[i0];
  const R b = Frac(x - dx[i0], dl, R());
#pragma unroll
  for (int i = 1; i < 3; ++i) {
    const R y1 = d[i0 + i], y2 = d[i1 + i];
    res[i - 1] = tmax((R)0.0, y1 + b * (y2 - y1));
  }
}
template <typename R>
HD R SplineLog4(int nd, const double* dx, const R* d, double x, R logx, R lmin, R ild, int iwhich) {  // :178-187
  const int idx = LogBin(logx, lmin, ild, nd);
  const int i0 = 9 * idx, i1 = i0 + 9;
  iwhich = (iwhich - 1) * 2 + 1;
  return Spline(dx[i0], dx[i1], d[i0 + iwhich], d[i1 + iwhich], d[i0 + iwhich + 1], d[i1 + iwhich + 1], x);
}
HD int FindLowerBin(const double* xd, int num, double x, int step) {  // FindLowerBinIndex :227-236
  int ml = -1, mu = num - 1;
  while (abs(mu - ml) > 1) {
    const int mav = (int)(0.5 * (ml + mu));
    if (x < xd[step * mav]) mu = mav; else ml = mav;
  }
  return mu - 1;
}
// FindLowerBin on a non-decreasing table, trying the bins hint-3 .. hint
// first (five independent loads instead of a chain of dependent ones). On such
// a table the search returns the largest k <= num-2 with xd[step*k] <= x
// whatever path it takes, so a bin that satisfies that is the same answer.
HD int FindLowerBinHint(const double* xd, int num, double x, int step, int hint) {
  if (hint >= 3 && hint <= num - 2) {
    const double v0 = xd[step * (hint - 3)], v1 = xd[step * (hint - 2)], v2 = xd[step * (hint - 1)];
    const double v3 = xd[step * hint], v4 = xd[step * (hint + 1)];
    if (v3 <= x && (x < v4 || hint == num - 2)) return hint;
    if (v2 <= x && x < v3) return hint - 1;
    if (v1 <= x && x < v2) return hint - 2;
    if (v0 <= x && x < v1) return hint - 3;
  }
  return FindLowerBin(xd, num, x, step);
}

// ---- track state held in registers (G4HepEmTrack, G4HepEmElectronTrack,
//      G4HepEmGammaTrack, G4HepEmMSCTrackData) ----------------------------------------------
template <typename R> struct Trk {
  R dir[3];
  double ekin;                   // kinetic energy stays double in both forms
  R lekin;
  double edep;
  R gstep, safety;
  R mfp[4], nial[4];
  int imc, winner;
  bool onBoundary, isEl;
  HD void setE(double e) { ekin = e; lekin = (R)100.0; }
  HD R logE() {                  // G4HepEmTrack::GetLogEKin (G4HepEmTrack.hh:95-100)
    if (lekin > (R)99.0) lekin = (ekin > 0.) ? vlog((R)ekin) : (R)-30;
    return lekin;
  }
};
template <typename R> struct Msc {
  R lambtr1, trueStep, zpath, displ[3], dir[3], initRange, dynRangeFactor, tlimitMin, par1, par2, par3;
  bool noScat, isDisplace, firstStep, active;
};
template <typename R> struct Sec { int ptype; double ekin; R dir[3]; };
// Secondaries of one step. Each interaction creates at most two, all of one
// kind (electron list or gamma list of G4HepEmTLData), so creation order is
// also the reference's stacking order (electrons first, then gammas) and the
// two slots are addressed statically (no run-time indexed arrays).
template <typename R> struct SecList {
  int ns;
  Sec<R> s0, s1;
};
// G4HepEmTrack::SetNumIALeft / GetMFP at a run-time process index, written as
// selects over the four slots so the arrays stay in registers
template <typename R> HD void ResetNumIALeft(R* nial, int ip) {
  // every slot is written with a select: a conditional store at a run-time
  // slot would be merged into one indexed store and put the array in scratch
  nial[0] = ip == 0 ? (R)-1.0 : nial[0];
  nial[1] = ip == 1 ? (R)-1.0 : nial[1];
  nial[2] = ip == 2 ? (R)-1.0 : nial[2];
  nial[3] = ip == 3 ? (R)-1.0 : nial[3];
}
template <typename R> HD R PickMFP(const R* mfp, int ip) {
  return ip == 0 ? mfp[0] : ip == 1 ? mfp[1] : ip == 2 ? mfp[2] : mfp[3];
}

// ---- G4HepEmElectronManager.icc: table accessors ------------------------------------------
// the [y, y''] column `col` (0 range, 1 dE/dx) of couple imc's loss data at ekin
template <typename R> HD R LossSpline(const DevEl<R>& e, int imc, int col, double ekin, R lekin) {
  return SplineLog2(e.N, e.lossGrid64, e.lossData + e.N * (5 * imc + 2 * col), ekin, lekin, e.lossLogMin, e.lossILD,
                    e.lossEmin, e.lossEmax);
}
template <typename R> HD R GetRestRange(const DevEl<R>& e, int imc, double ekin, R lekin) {  // :486-493
  return tmax((R)0.0, LossSpline(e, imc, 0, ekin, lekin));
}
template <typename R> HD R GetRestDEDX(const DevEl<R>& e, int imc, double ekin, R lekin) {  // :495-502
  return tmax((R)0.0, LossSpline(e, imc, 1, ekin, lekin));
}
// :504-520; the low-range quadratic branch is an FP64 island in both forms
// binHint: the loss-grid bin of an energy whose range is near `range` (at or
// above it), or -1
template <typename R> HD double GetInvRange(const DevEl<R>& e, int imc, double range, int binHint = -1) {
  const int N = e.N;
  const double* xd = e.lossData64 + 5 * N * imc;
  const double minRange = xd[0];
  if (range < minRange) {
    const double dum = range / minRange;
    return tmax(0.0, e.lossGrid64[0] * dum * dum);
  }
  const int i = e.rangeMonotone ? FindLowerBinHint(xd, N, range, 2, binHint) : FindLowerBin(xd, N, range, 2);
  const R* sd = e.lossData + 5 * N * imc + 4 * N;
  return (double)tmax((R)0.0, Spline(xd[2 * i], xd[2 * (i + 1)], e.lossGrid[i], e.lossGrid[i + 1], sd[i], sd[i + 1], range));
}
// GetInvRange of the range of an energy in loss-grid bin `bin`, whose grid
// energies eLo, eHi the caller holds: on a non-decreasing range column the
// search ends in that bin whenever range lies in [R_bin, R_bin+1) (or beyond
// R_bin in the last bin), and the spline then takes eLo, eHi as its ordinates;
// otherwise the general search.
template <typename R> HD double GetInvRangeInBin(const DevEl<R>& e, int imc, double range, int bin, double eLo, double eHi) {
  const int N = e.N;
  const double* xd = e.lossData64 + 5 * N * imc;
  if (e.rangeMonotone && range >= xd[0]) {
    const double r1 = xd[2 * bin], r2 = xd[2 * (bin + 1)];
    if (r1 <= range && (range < r2 || bin == N - 2)) {
      const R* sd = e.lossData + 5 * N * imc + 4 * N;
      return (double)tmax((R)0.0, Spline(r1, r2, (R)eLo, (R)eHi, sd[bin], sd[bin + 1], range));
    }
  }
  return GetInvRange(e, imc, range, bin);
}
template <typename R> HD R GetRestMacXSec(const DevEl<R>& e, int imc, double ekin, R lekin, bool isioni) {  // :522-532
  const XsBlock<R> b = e.xsBlock[2 * imc + (isioni ? 0 : 1)];
  if (ekin < b.minE) return (R)0.0;
  return tmax((R)0.0,
              SplineLog3(b.n, e.xsData64 + b.start, e.xsData + b.start, ekin, lekin, b.logMin, b.ild, b.minE, b.lastE));
}
template <typename R> HD R GetMacXSecNuclear(const DevEl<R>& e, int imat, double ekin, R lekin) {  // :534-541
  if (ekin < e.nucEmin) return (R)0.0;
  return tmax((R)0.0, SplineLog2(e.nucN, e.nucGrid64, e.nucData + imat * 2 * e.nucN, ekin, lekin, e.nucLogMin, e.nucILD,
                                 e.nucEmin, e.nucEmax));
}
template <typename R> HD R GetRestMacXSecForStepping(const DevEl<R>& e, int imc, double ekin, R lekin, bool isioni) {  // :544-568
  const R log08 = (R)-0.22314355131420971;
  const XsBlock<R> b = e.xsBlock[2 * imc + (isioni ? 0 : 1)];
  if (ekin > b.maxE) {
    const double ekinReduced = 0.8 * ekin;
    if (ekinReduced < b.maxE) return tmax((R)0.0, b.maxV);
    ekin = ekinReduced;
    lekin += log08;
  }
  if (ekin < b.minE) return (R)0.0;
  return tmax((R)0.0,
              SplineLog3(b.n, e.xsData64 + b.start, e.xsData + b.start, ekin, lekin, b.logMin, b.ild, b.minE, b.lastE));
}
template <typename R> HD R GetTransportMFP(const DevEl<R>& e, int im, double ekin, R lekin) {  // :576-583
  const R tr = tmax((R)0.0, SplineLog2(e.N, e.lossGrid64, e.tr1 + 2 * e.N * im, ekin, lekin, e.lossLogMin, e.lossILD,
                                       e.lossEmin, e.lossEmax));
  return tr > (R)0. ? (R)1. / tr : (R)kALargeValue;
}
template <typename R> HD R MacXsecAnnihilation(R ekin, R eDens) {  // :585-593
  const R tau = ekin * (R)kInvElectronMassC2;
  const R gam = tau + (R)1.0;
  const R gam2 = gam * gam;
  const R bg2 = tau * (tau + (R)2.0);
  const R bg = dsqrt(bg2);
  return eDens * (R)kPir02 * ((gam2 + (R)4. * gam + (R)1.) * vlog(gam + bg) - (gam + (R)3.) * bg) / (bg2 * (gam + (R)1.));
}

// ---- G4HepEmElectronInteractionUMSC.icc ----------------------------------------------------
template <typename R>
HD void UMSCStepLimit(const DevTables<R>& T, Msc<R>& m, R ekin, int imat, int ireg, R range, R presafety, bool onBoundary,
                      bool isEl, Rng& rng) {  // :19-126
  m.noScat = false;
  m.isDisplace = true;
  const R kTLimitMinfix = (R)1.0E-8;
  if (m.trueStep < kTLimitMinfix || range * (R)T.matUMSCPar[imat] < presafety) {
    m.isDisplace = false;
    return;
  }
  const R rangeFactor = (R)T.regRangeFac[ireg], safetyFactor = (R)T.regSafetyFac[ireg];
  R tlimit;
  if (!T.regMinimal[ireg]) {
    if (m.firstStep || onBoundary) {
      const R lambdaTr1 = m.lambtr1;
      m.initRange = tmax(range, lambdaTr1);
      m.dynRangeFactor = lambdaTr1 > (R)1.0 ? rangeFactor * ((R)0.75 + (R)0.25 * lambdaTr1) : rangeFactor;
      const R stepMin = lambdaTr1 * (R)1.0E-3 /
                        ((R)2.0E-3 + ekin * ((R)T.matStepMin[2 * imat] + ekin * (R)T.matStepMin[2 * imat + 1]));
      const R dum0 = isEl ? (R)0.87 * (R)T.matZeff23[imat] : (R)0.70 * (R)T.matZeffSqrt[imat];
      const R dum1 = ekin > (R)5.0E-3 ? dum0 * stepMin : dum0 * stepMin * (R)0.5 * ((R)1.0 + ekin * (R)200.0);
      m.tlimitMin = tmax(dum1, kTLimitMinfix);
      m.firstStep = false;
    }
    const R tlimitmin = m.tlimitMin;
    tlimit = range > presafety ? tmax(tmax(m.initRange * m.dynRangeFactor, safetyFactor * presafety), tlimitmin)
                               : tmax(range, tlimitmin);
  } else {
    if (onBoundary) {
      const R lambdaTr1 = m.lambtr1;
      const R tmp = range > lambdaTr1 ? rangeFactor * range : rangeFactor * lambdaTr1;
      m.initRange = tmax(tmp, (R)10 * kTLimitMinfix);
    }
    tlimit = m.initRange;
  }
  const R tlimitmin = m.tlimitMin;
  if (tlimit < m.trueStep) {
    const R dum0 = tlimit > tlimitmin ? tmax(rng.Gauss(tlimit, (R)0.1 * (tlimit - tlimitmin)), tlimitmin) : tlimitmin;
    m.trueStep = tmin(dum0, m.trueStep);
  }
}

template <typename R> HD R UMSCSimpleScattering(R xmeanth, R x2meanth, Rng& rng) {  // :280-293
  const R dum0 = (R)3. * x2meanth - (R)1;
  const R dum1 = (R)2. * xmeanth - dum0;
  const R a = (R)1. + (R)4. * dum0 / dum1;
  const R prob = ((R)2. + a) * xmeanth / a;
  R r[2];
  rng.flatArray(2, r);
  return (r[0] < prob) ? (R)-1. + (R)2. * vpow(r[1], (R)1. / ((R)1. + a)) : (R)-1. + (R)2. * r[1];
}

template <typename R> HD R UMSCTheta0PositronCorrection(R eekin, R zeff) {  // :313-338
  const R ff = (R)1. + zeff * ((R)1.84035E-4 * zeff - (R)1.86427E-2) + (R)0.41125;
  const R a = (R)0.994 - (R)4.08E-3 * zeff;
  const R b = (R)7.16 + ((R)52.6 + (R)365. / zeff) / zeff;
  const R tu = dsqrt(eekin) * (R)kInvElectronMassC2;
  const R x = dsqrt(tu * (tu + (R)2.) / ((tu + (R)1.) * (tu + (R)1.)));
  const R xl = (R)0.6;
  if (x < xl) return ff * a * ((R)1. - vexp(-b * x));
  const R c = (R)1.00 - (R)4.47E-3 * zeff;
  const R dd = (R)1.21E-3 * zeff;
  const R e = (R)113.0;
  const R xh = (R)0.9;
  if (x > xh) return ff * (c + dd * vexp(e * (x - (R)1.)));
  const R yl = a * ((R)1. - vexp(-b * xl));
  const R yh = c + dd * vexp(e * (xh - (R)1.));
  const R y0 = (yh - yl) / (xh - xl);
  const R y1 = yl - y0 * xl;
  return ff * (y0 * x + y1);
}

template <typename R>
HD R UMSCComputeTheta0(R stepInRadLength, R postEkin, R preEkin, R zeff, R tc0, R tc1, bool isEl, bool isPosCor) {  // :297-309
  const R kHighland = (R)13.6;
  const R me = (R)kElectronMassC2;
  const R postInvBetaPc = (postEkin + me) / (postEkin * (postEkin + (R)2. * me));
  const R invBetaPc = preEkin != postEkin ? dsqrt(postInvBetaPc * (preEkin + me) / (preEkin * (preEkin + (R)2. * me))) : postInvBetaPc;
  const R y = (isEl || !isPosCor) ? stepInRadLength : stepInRadLength * UMSCTheta0PositronCorrection(preEkin * postEkin, zeff);
  return kHighland * dsqrt(y) * invBetaPc * (tc0 + tc1 * vlog(y));
}

template <typename R>
HD R UMSCSampleCosineTheta(const DevTables<R>& T, int imat, R pStep, R preEkin, R preTr1mfp, R postEkin, R postTr1mfp,
                           R tlimitMin, bool isEl, bool isPosCor, Rng& rng) {  // :152-277
  const R radLength = (R)T.matRadLen[imat], zeff = (R)T.matZeff[imat];
  const double* tail = T.matTail + 4 * imat;
  const double* theta = T.matTheta + 2 * imat;
  const R iPreTr1mfp = (R)1.0 / preTr1mfp;
  const R deltaR1mfp = preTr1mfp - postTr1mfp;
  const R tau = dabs(deltaR1mfp) > (R)0.01 * preTr1mfp ? pStep * vlog(preTr1mfp / postTr1mfp) / deltaR1mfp : pStep * iPreTr1mfp;
  if (tau > (R)8.0) return (R)2.0 * (R)rng.flat() - (R)1.0;
  const R kTauSmall = (R)1.0E-16;
  if (tau < kTauSmall) return (R)1.0;
  R xmeanth, x2meanth;
  if (tau < (R)0.01) {
    xmeanth = (R)1.0 - tau * ((R)1.0 - (R)0.5 * tau);
    x2meanth = (R)1.0 - tau * ((R)5.0 - (R)6.25 * tau) * (R)0.333333;
  } else {
    xmeanth = vexp(-tau);
    x2meanth = ((R)1.0 + (R)2.0 * vexp((R)-2.5 * tau)) * (R)0.333333;
  }
  if (postEkin < (R)0.5 * preEkin) return UMSCSimpleScattering(xmeanth, x2meanth, rng);
  const R tsmall = tmin(tlimitMin, (R)1.0);
  const bool notSmall = pStep > tsmall;
  const R theta0 = notSmall
      ? UMSCComputeTheta0(pStep / radLength, postEkin, preEkin, zeff, (R)theta[0], (R)theta[1], isEl, isPosCor)
      : UMSCComputeTheta0(tsmall / radLength, postEkin, preEkin, zeff, (R)theta[0], (R)theta[1], isEl, isPosCor) * dsqrt(pStep / tsmall);
  if (theta0 > (R)kPi * (R)0.166666) return UMSCSimpleScattering(xmeanth, x2meanth, rng);
  const R theta2 = theta0 * theta0;
  if (theta2 < kTauSmall) return (R)1.0;
  const R dumtau = notSmall ? tau : tsmall * iPreTr1mfp;
  const R parU = vpow(dumtau, (R)0.1666666);
  const R dumxsi = (R)tail[0] + parU * ((R)tail[1] + parU * (R)tail[2]) + (R)tail[3] * vlog(pStep / (tau * radLength));
  const R parXsi = tmax(dumxsi, (R)1.9);
  const R parC = dabs(parXsi - (R)3.) < (R)0.001 ? (R)3.001 : dabs(parXsi - (R)2.) < (R)0.001 ? (R)2.001 : parXsi;
  const R dumC1 = parC - (R)1.;
  const R dumEa = vexp(-parXsi);
  const R dumEaa = (R)1. / ((R)1. - dumEa);
  R thex = theta2 * ((R)1.0 - theta2 * (R)0.0833333);
  if (theta2 > (R)0.01) {
    const R dum = (R)2.0 * dsin((R)0.5 * theta0);
    thex = dum * dum;
  }
  const R xmean1 = (R)1. - ((R)1. - ((R)1. + parXsi) * dumEa) * thex * dumEaa;
  if (xmean1 <= (R)0.999 * xmeanth) return UMSCSimpleScattering(xmeanth, x2meanth, rng);
  const R x0 = (R)1. - parXsi * thex;
  const R bx = parC * thex;
  const R b = bx + x0;
  const R b1 = b + (R)1.;
  const R eb1 = vpow(b1, dumC1);
  const R ebx = vpow(bx, dumC1);
  const R dd = ebx / eb1;
  const R xmean2 = (x0 + dd - (bx - b1 * dd) / (parC - (R)2.)) / ((R)1. - dd);
  const R f1x0 = dumEa * dumEaa;
  const R f2x0 = dumC1 / (parC * ((R)1. - dd));
  const R prob = f2x0 / (f1x0 + f2x0);
  const R qprb = xmeanth / (prob * xmean1 + ((R)1. - prob) * xmean2);
  R r[3];
  rng.flatArray(3, r);
  if (r[0] < qprb) {
    if (r[1] < prob) return (R)1. + vlog(dumEa + r[2] / dumEaa) * thex;
    const R var0 = ((R)1.0 - dd) * r[2];
    if (var0 < (R)0.01 * dd) {
      const R var = var0 / (dd * dumC1);
      return (R)-1.0 + var * ((R)1.0 - var * (R)0.5 * parC) * b1;
    }
    return (R)1.0 + thex * (parC - parXsi - parC * vpow(var0 + dd, (R)-1. / dumC1));
  }
  return (R)2.0 * r[1] - (R)1.0;
}

template <typename R> HD void UMSCSampleDisplacement(R pStep, R thePhi, Msc<R>& m, Rng& rng) {  // :342-357
  const R r = (R)0.73 * dsqrt((pStep - m.zpath) * (pStep + m.zpath));
  const R cbeta = (R)2.16;
  const R cbeta1 = (R)1. - vexp(-cbeta * (R)kPi);
  R rr[2];
  rng.flatArray(2, rr);
  const R psi = -vlog((R)1. - rr[0] * cbeta1) / cbeta;
  const R phi = (rr[1] < (R)0.5) ? thePhi + psi : thePhi - psi;
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  m.displ[0] = r * cosPhi;
  m.displ[1] = r * sinPhi;
  m.displ[2] = (R)0.0;
}

template <typename R>
HD void UMSCSampleScattering(const DevTables<R>& T, Msc<R>& m, R pStep, R preEkin, R preTr1mfp, R postEkin, R postTr1mfp, int imat,
                             bool isEl, bool isPosCor, bool isDisplace, Rng& rng) {  // :129-149
  const R cost = UMSCSampleCosineTheta(T, imat, pStep, preEkin, preTr1mfp, postEkin, postTr1mfp, m.tlimitMin, isEl, isPosCor, rng);
  if (dabs(cost) >= (R)1.0) {
    m.noScat = true;
    return;
  }
  const R sth = dsqrt(((R)1.0 - cost) * ((R)1.0 + cost));
  const R phi = (R)k2Pi * (R)rng.flat();
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  m.dir[0] = sth * cosPhi;
  m.dir[1] = sth * sinPhi;
  m.dir[2] = cost;
  m.isDisplace = m.isDisplace && isDisplace;
  if (m.isDisplace && pStep > m.zpath) UMSCSampleDisplacement(pStep, phi, m, rng);
}

// ---- G4HepEmElectronEnergyLossFluctuation.icc -------------------------------------------------
template <typename R> HD R SampleGaussianLoss(R meane, R sig2e, Rng& rng) {  // :100-111
  const R twom = (R)2. * meane;
  if (meane * meane < (R)0.0625 * sig2e) return twom * (R)rng.flat();
  const R sig = dsqrt(sig2e);
  R eloss;
  do {
    eloss = rng.Gauss(meane, sig);
  } while (eloss < (R)0. || eloss > twom);
  return eloss;
}

template <typename R> HD R SampleELossFluctuation(R tcut, R excEner, R meanELoss, Rng& rng) {  // :11-97
  const R scaling = tmin((R)1. + (R)5.E-4 / tcut, (R)1.5);
  const R meanLoss = meanELoss / scaling;
  const R kRate = (R)0.56, kE0 = (R)1.E-5, kNMaxCont = (R)8.;
  const R w1 = tcut / kE0;
  R a3 = meanLoss * (tcut - kE0) / (kE0 * tcut * vlog(w1));
  R a1 = (R)0., e1 = excEner, eloss = (R)0.0;
  if (tcut > excEner) {
    const R a1Tmp = meanLoss * ((R)1. - kRate) / excEner;
    const R kA0 = (R)42., kFw = (R)4.;
    const R dum0 = a1Tmp < kA0 ? (R).1 + (kFw - (R).1) * dsqrt(a1Tmp / kA0) : kFw;
    a1 = a1Tmp / dum0;
    e1 *= dum0;
    a3 *= kRate;
    if (a1 > kNMaxCont) {
      const R emean = a1 * e1;
      const R sig2e = emean * e1;
      eloss = SampleGaussianLoss(emean, sig2e, rng);
    } else {
      const int pp = rng.Poisson(a1);
      eloss = pp > 0 ? ((R)(pp + 1) - (R)2. * (R)rng.flat()) * e1 : (R)0.;
    }
  }
  if (a3 > (R)0.) {
    R p3 = a3, alfa = (R)1.;
    if (a3 > kNMaxCont) {
      alfa = w1 * (kNMaxCont + a3) / (w1 * kNMaxCont + a3);
      const R alfa1 = alfa * vlog(alfa) / (alfa - (R)1.);
      const R namean = a3 * w1 * (alfa - (R)1.) / ((w1 - (R)1.) * alfa);
      const R emean = namean * kE0 * alfa1;
      const R sig2e = kE0 * kE0 * namean * (alfa - alfa1 * alfa1);
      eloss += SampleGaussianLoss(emean, sig2e, rng);
      p3 = a3 - namean;
    }
    const R w3 = alfa * kE0;
    if (tcut > w3) {
      const R w = (tcut - w3) / tcut;
      const int nnb = rng.Poisson(p3);
      // the reference draws these in blocks of 8 (flatArray); the stream is the
      // same sequence of single draws, summed in the same order
      for (int i = 0; i < nnb; ++i) eloss += w3 / ((R)1. - w * (R)rng.flat());
    }
  }
  return eloss * scaling;
}

// ---- G4HepEmInteractionUtils.{hh,icc} -------------------------------------------------------
// LPM G(s), Phi(s) tabulated on s in [0, 2], step 0.05 (Stanev et al. approximation).
__constant__ double kFuncLPMd[84] = {
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

template <typename R> HD R SampleCostModifiedTsai(R primEkin, Rng& rng) {  // .icc:11-25
  const R uMax = (R)2.0 * ((R)1.0 + primEkin * (R)kInvElectronMassC2);
  R r[3], u;
  do {
    rng.flatArray(3, r);
    const R uu = -vlog(r[0] * r[1]);
    u = ((R)0.25 > r[2]) ? uu * (R)1.6 : uu * (R)0.533333333;
  } while (u > uMax);
  return (R)1.0 - (R)2.0 * u * u / (uMax * uMax);
}

template <typename R>
HD void EvaluateLPMFunctions(R& funcXiS, R& funcGS, R& funcPhiS, R egamma, R etotal, R elpm, R z23, R ilVarS1, R ilVarS1Cond,
                             R densityCor, R times) {  // .icc:28-76
  const R sqrt2 = (R)1.414213562373095;
  const R redegamma = egamma / etotal;
  const R varSprime = dsqrt((R)0.125 * redegamma * elpm / (times * ((R)1.0 - redegamma) * etotal));
  const R varS1 = z23 / ((R)184.15 * (R)184.15);
  const R condition = sqrt2 * varS1;
  R funcXiSprime = (R)2.0;
  if (varSprime > (R)1.0) {
    funcXiSprime = (R)1.0;
  } else if (varSprime > condition) {
    const R funcHSprime = vlog(varSprime) * ilVarS1Cond;
    funcXiSprime = (R)1.0 + funcHSprime - (R)0.08 * ((R)1.0 - funcHSprime) * funcHSprime * ((R)2.0 - funcHSprime) * ilVarS1Cond;
  }
  funcXiS = funcXiSprime;
  const R varS = varSprime / dsqrt(funcXiSprime);
  R varShat = varS;
  if (densityCor != (R)0.0) {
    varShat *= ((R)1.0 + densityCor / (egamma * egamma));
    funcXiS = (R)2.0;
    if (varShat > (R)1.0) funcXiS = (R)1.0;
    else if (varShat > varS1) funcXiS = (R)1.0 + vlog(varShat) * ilVarS1;
  }
  if (varShat < (R)2.0) {
    R val = varShat * (R)20.0;
    int ilow = (int)val;
    val -= (R)ilow;
    ilow *= 2;
    funcGS = ((R)kFuncLPMd[ilow + 2] - (R)kFuncLPMd[ilow]) * val + (R)kFuncLPMd[ilow];
    funcPhiS = ((R)kFuncLPMd[ilow + 3] - (R)kFuncLPMd[ilow + 1]) * val + (R)kFuncLPMd[ilow + 1];
  } else {
    R ss = (R)1.0 / (varShat * varShat);
    ss *= ss;
    funcGS = (R)1.0 - (R)0.0230655 * ss;
    funcPhiS = (R)1.0 - (R)0.01190476 * ss;
  }
  if (funcXiS * funcPhiS > (R)1.0 || varShat > (R)0.57) funcXiS = (R)1.0 / funcPhiS;
}

// ---- G4HepEmElectronInteractionIoni.icc ---------------------------------------------------------
template <typename R> HD R SampleETransferMoller(R elCut, R primEkin, Rng& rng) {  // :50-74
  const R tmin_ = elCut, tmax_ = (R)0.5 * primEkin;
  const R xmin = tmin_ / primEkin, xmax = tmax_ / primEkin;
  const R gamma = primEkin * (R)kInvElectronMassC2 + (R)1.0;
  const R gamma2 = gamma * gamma;
  const R xminmax = xmin * xmax;
  const R gg = ((R)2.0 * gamma - (R)1.0) / gamma2;
  const R y = (R)1. - xmax;
  const R gf = (R)1.0 - gg * xmax + xmax * xmax * ((R)1.0 - gg + ((R)1.0 - gg * y) / (y * y));
  R dum, r[2], delta;
  do {
    rng.flatArray(2, r);
    delta = xminmax / (xmin * ((R)1.0 - r[0]) + xmax * r[0]);
    const R xx = (R)1.0 - delta;
    dum = (R)1.0 - gg * delta + delta * delta * ((R)1.0 - gg + ((R)1.0 - gg * xx) / (xx * xx));
  } while (gf * r[1] > dum);
  return delta * primEkin;
}
template <typename R> HD R SampleETransferBhabha(R elCut, R primEkin, Rng& rng) {  // :76-108
  const R tmin_ = elCut, tmax_ = primEkin;
  const R xmin = tmin_ / primEkin, xmax = tmax_ / primEkin;
  const R gamma = primEkin * (R)kInvElectronMassC2 + (R)1.0;
  const R gamma2 = gamma * gamma;
  const R beta2 = (R)1. - (R)1. / gamma2;
  const R xminmax = xmin * xmax;
  const R y = (R)1.0 / ((R)1.0 + gamma), y2 = y * y, y12 = (R)1.0 - (R)2.0 * y;
  const R b1 = (R)2.0 - y2, b2 = y12 * ((R)3.0 + y2), y122 = y12 * y12, b4 = y122 * y12, b3 = b4 + y122;
  const R xmax2 = xmax * xmax;
  const R gf = (R)1.0 + (xmax2 * b4 - xmin * xmin * xmin * b3 + xmax2 * b2 - xmin * b1) * beta2;
  R dum, r[2], delta;
  do {
    rng.flatArray(2, r);
    delta = xminmax / (xmin * ((R)1.0 - r[0]) + xmax * r[0]);
    const R xx = delta * delta;
    dum = (R)1.0 + (xx * xx * b4 - delta * xx * b3 + xx * b2 - delta * b1) * beta2;
  } while (gf * r[1] > dum);
  return delta * primEkin;
}
template <typename R> HD void IoniSampleDirections(R primEkin, R deltaEkin, R* secDir, R* primDir, Rng& rng) {  // :111-138
  const R me = (R)kElectronMassC2;
  const R elInitETot = primEkin + me;
  const R elInitPTot = dsqrt(primEkin * (elInitETot + me));
  const R deltaPTot = dsqrt(deltaEkin * (deltaEkin + (R)2.0 * me));
  const R cost = deltaEkin * (elInitETot + me) / (deltaPTot * elInitPTot);
  const R cosTheta = tmax((R)-1.0, tmin(cost, (R)1.0));
  const R sinTheta = dsqrt(((R)1.0 - cosTheta) * ((R)1.0 + cosTheta));
  const R phi = (R)k2Pi * (R)rng.flat();
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  secDir[0] = sinTheta * cosPhi;
  secDir[1] = sinTheta * sinPhi;
  secDir[2] = cosTheta;
  Rotate(secDir, primDir);
  primDir[0] = elInitPTot * primDir[0] - deltaPTot * secDir[0];
  primDir[1] = elInitPTot * primDir[1] - deltaPTot * secDir[1];
  primDir[2] = elInitPTot * primDir[2] - deltaPTot * secDir[2];
  const R norm = (R)1.0 / dsqrt(primDir[0] * primDir[0] + primDir[1] * primDir[1] + primDir[2] * primDir[2]);
  primDir[0] *= norm;
  primDir[1] *= norm;
  primDir[2] *= norm;
}
template <typename R> HD void IoniPerform(const DevTables<R>& T, Trk<R>& t, SecList<R>& sl, Rng& rng) {  // :19-47
  const double primEkin = t.ekin;
  const R elCut = (R)T.mcElCut[t.imc];
  const R pE = (R)primEkin;
  const R maxET = t.isEl ? (R)0.5 * pE : pE;
  if (maxET <= elCut) return;
  const R delta = t.isEl ? SampleETransferMoller(elCut, pE, rng) : SampleETransferBhabha(elCut, pE, rng);
  Sec<R>* s = &sl.s0;
  sl.ns = 1;
  s->ptype = 0;
  IoniSampleDirections(pE, delta, s->dir, t.dir, rng);
  t.setE(primEkin - (double)delta);
  s->ekin = (double)delta;
}

// ---- G4HepEmElectronInteractionBrem.icc ---------------------------------------------------------
template <typename R>
HD int BremSelectTargetAtom(const DevEl<R>& e, int imc, double ekin, R lekin, R urndn, bool isSB) {  // :266-297
  const int start = isSB ? e.selSBStart[imc] : e.selRBStart[imc];
  const double* d64 = (isSB ? e.selSB64 : e.selRB64) + start;
  const R* data = (isSB ? e.selSB : e.selRB) + start;
  const int numData = (int)d64[0];
  const int numElem = (int)d64[1];
  const R logE0 = data[2], invLD = data[3];
  const double* xd64 = d64 + 4;
  const R* xd = data + 4;
  const double xv = tmax(xd64[0], tmin(xd64[numElem * (numData - 1)], ekin));
  const int idxEkin = (int)tmax((R)0.0, tmin((lekin - logE0) * invLD, (R)(numData - 2)));
  int i0 = idxEkin * numElem, i1 = i0 + numElem;
  const double x1 = xd64[i0++], x2 = xd64[i1++];
  const double dl = x2 - x1;
  const R b = Frac(xv - x1, dl, R());
  int ie = 0;
  while (ie < numElem - 1 && urndn > xd[i0 + ie] + b * (xd[i1 + ie] - xd[i0 + ie])) ++ie;
  return ie;
}
// First i (a multiple of 3, i < 3*size) with v[i] > val, or 3*size: the
// reference's forward scan (:328-344) started at i0, which must satisfy
// v[j] <= val for every j < i0 so that the result is the scan's from 0.
template <typename R> HD int BremLinSearch(const R* v, int size, R val, int i0 = 0) {
  int i = i0;
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
template <typename R>
HD R SampleETransferSB(const DevTables<R>& T, double primEkinD, R lEkin, int imc, Rng& rng, bool isEl) {  // :73-182
  const R primEkin = (R)primEkinD;
  const R gamCut = (R)T.mcGamCut[imc], logGamCut = (R)T.mcLogGamCut[imc];
  const int imat = T.mcMat[imc];
  const int nel = T.matNElem[imat];
  const int elemIndx = nel > 1 ? BremSelectTargetAtom(T.el[isEl ? 0 : 1], imc, primEkinD, lEkin, (R)rng.flat(), true) : 0;
  const int iZet = T.matElemZ[T.matElemStart[imat] + elemIndx];
  const R dZet = (R)iZet;
  const int iStart = T.sbStartZ[iZet];
  const int iGamCut = T.sbGCutIdx[T.sbGCutStart[imc] + elemIndx];
  bool isCorner = false, isSimply = false;
  int eIdx = (int)T.sbData[iStart + 2];
  if (primEkin < (R)T.sbElE[eIdx]) {
    const R val = (lEkin - (R)T.sbLogMinE) * (R)T.sbILDE;
    eIdx = (int)val;
    R pIndxH = val - (R)eIdx;
    if ((R)T.sbElE[eIdx] <= gamCut) {
      pIndxH = (lEkin - logGamCut) / ((R)T.sbLElE[eIdx + 1] - logGamCut);
      isCorner = true;
    }
    if ((R)rng.flat() < pIndxH) ++eIdx;
    else if (isCorner) isSimply = true;
  }
  const int minEIndx = (int)T.sbData[iStart + 1];
  const int numGamCuts = (int)T.sbData[iStart + 3];
  const int sizeOneE = numGamCuts + 3 * T.sbNumKappa;
  const int iSTStart = iStart + 4 + (eIdx - minEIndx) * sizeOneE;
  const R minV = T.sbData[iSTStart + iGamCut];
  const R* stData = T.sbData + iSTStart + numGamCuts;
  const uint8_t* guide = T.sbGuide + (size_t)(T.sbGuideStartZ[iZet] + (eIdx - minEIndx)) * kSBGuide;
  const R lKTrans = (logGamCut - lEkin) / (logGamCut - (R)T.sbLElE[eIdx]);
  const R primETot = primEkin + (R)kElectronMassC2;
  const R dielSupConst = (R)T.matDensCor[imat] * primETot * primETot;
  R suppression, r[2], eGamma;
  do {
    rng.flatArray(2, r);
    R kappa;
    if (!isSimply) {
      const R cumRV = r[0] * ((R)1.0 - minV) + minV;
      // start one guide bucket below the sample: every cumulative value
      // before the start is <= cumRV, so the scan finds the same index
      const int g = tmax(0, tmin(kSBGuide - 1, (int)(cumRV * (R)kSBGuide) - 1));
      const int cumLIndx3 = BremLinSearch(stData, T.sbNumKappa, cumRV, 3 * (int)guide[g]) - 3;
      const int cumLIndx = cumLIndx3 / 3;
      const R cumL = stData[cumLIndx3];
      const R pA = stData[cumLIndx3 + 1];
      const R pB = stData[cumLIndx3 + 2];
      const R cumH = stData[cumLIndx3 + 3];
      const R lKL = T.sbLKappa[cumLIndx];
      const R lKH = T.sbLKappa[cumLIndx + 1];
      const R dm1 = (cumRV - cumL) / (cumH - cumL);
      const R dm2 = ((R)1.0 + pA + pB) * dm1;
      const R dm3 = (R)1.0 + dm1 * (pA + pB * dm1);
      const R lKappa = lKL + dm2 / dm3 * (lKH - lKL);
      kappa = vexp(lKappa * lKTrans);
    } else {
      kappa = (R)1.0 - r[0] * ((R)1.0 - gamCut / primEkin);
    }
    eGamma = kappa * primEkin;
    const R invEGamma = (R)1.0 / eGamma;
    suppression = (R)1.0 / ((R)1.0 + dielSupConst * invEGamma * invEGamma);
    if (!isEl) {
      const R me = (R)kElectronMassC2;
      const R e1 = primEkin - gamCut;
      const R iBeta1 = (e1 + me) / dsqrt(e1 * (e1 + (R)2.0 * me));
      const R e2 = primEkin - eGamma;
      const R iBeta2 = (e2 + me) / dsqrt(e2 * (e2 + (R)2.0 * me));
      const R dum = (R)kAlpha * (R)k2Pi * dZet * (iBeta1 - iBeta2);
      suppression = (dum > (R)-12.) ? suppression * vexp(dum) : (R)0.;
    }
  } while (r[1] > suppression);
  return eGamma;
}
template <typename R>
HD R SampleETransferRB(const DevTables<R>& T, double primEkinD, R lEkin, int imc, Rng& rng, bool isEl) {  // :184-263
  const R primEkin = (R)primEkinD;
  const R gamCut = (R)T.mcGamCut[imc];
  const int imat = T.mcMat[imc];
  const int nel = T.matNElem[imat];
  const int elemIndx = nel > 1 ? BremSelectTargetAtom(T.el[isEl ? 0 : 1], imc, primEkinD, lEkin, (R)rng.flat(), false) : 0;
  const int iZet = T.matElemZ[T.matElemStart[imat] + elemIndx];
  const R dZet = (R)iZet;
  const int z = iZet < T.maxZ ? iZet : T.maxZ;
  const R densityFactor = (R)kMigdalConst * (R)T.matElDens[imat];
  const R lpmEnergy = (R)kLPMconstant * (R)T.matRadLen[imat];
  const R lpmEnergyLim = dsqrt(densityFactor) * lpmEnergy;
  const R primTotalE = primEkin + (R)kElectronMassC2;
  const R densityCorr = densityFactor * primTotalE * primTotalE;
  const bool isLPMActive = primTotalE > lpmEnergyLim;
  const R zFactor1 = (R)T.elZF1[z];
  const R zFactor2 = ((R)1. + (R)1. / dZet) / (R)12.;
  const R rejFuncMax = zFactor1 + zFactor2;
  const R xmin = vlog(gamCut * gamCut + densityCorr);
  const R xrange = vlog(primEkin * primEkin + densityCorr) - xmin;
  R r[2], eGamma, funcVal;
  do {
    rng.flatArray(2, r);
    eGamma = dsqrt(tmax(vexp(xmin + r[0] * xrange) - densityCorr, (R)0.0));
    const R y = eGamma / primTotalE;
    const R onemy = (R)1. - y;
    const R dum0 = (R)0.25 * y * y;
    if (isLPMActive) {
      R funcGS, funcPhiS, funcXiS;
      EvaluateLPMFunctions(funcXiS, funcGS, funcPhiS, eGamma, primTotalE, lpmEnergy, (R)T.elZ23[z], (R)T.elILS1[z],
                           (R)T.elILS1Cond[z], densityCorr, (R)1.0);
      const R term1 = funcXiS * (dum0 * funcGS + (onemy + (R)2.0 * dum0) * funcPhiS);
      funcVal = term1 * zFactor1 + onemy * zFactor2;
    } else {
      const R dum1 = onemy + (R)3. * dum0;
      if (iZet < 5) {
        funcVal = dum1 * zFactor1 + onemy * zFactor2;
      } else {
        const R z13 = (R)T.elZ13[z];
        const R dum2 = y / (primTotalE - eGamma);
        const R gam = dum2 * (R)100. * (R)kElectronMassC2 / z13;
        const R eps = gam / z13;
        const R gam2 = gam * gam;
        const R phi1 = (R)16.863 - (R)2.0 * vlog((R)1.0 + (R)0.311877 * gam2) + (R)2.4 * vexp((R)-0.9 * gam) + (R)1.6 * vexp((R)-1.5 * gam);
        const R phi2 = (R)2.0 / ((R)3.0 + (R)19.5 * gam + (R)18.0 * gam2);
        const R eps2 = eps * eps;
        const R psi1 = (R)24.34 - (R)2.0 * vlog((R)1.0 + (R)13.111641 * eps2) + (R)2.8 * vexp((R)-8.0 * eps) + (R)1.2 * vexp((R)-29.2 * eps);
        const R psi2 = (R)2.0 / ((R)3.0 + (R)120.0 * eps + (R)1200.0 * eps2);
        const R logZ = (R)T.elLogZ[z];
        const R Fz = logZ / (R)3. + (R)T.elCoulomb[z];
        const R invZ = (R)1. / dZet;
        funcVal = dum1 * (((R)0.25 * phi1 - Fz) + ((R)0.25 * psi1 - (R)2. * logZ / (R)3.) * invZ) + (R)0.125 * onemy * (phi2 + psi2 * invZ);
      }
    }
    funcVal = tmax((R)0.0, funcVal);
  } while (funcVal < rejFuncMax * r[1]);
  return eGamma;
}
template <typename R> HD void BremSampleDirections(R primEkin, R secEkin, R* secDir, R* primDir, Rng& rng) {  // :299-322
  const R cost = SampleCostModifiedTsai(primEkin, rng);
  const R sint = dsqrt(((R)1.0 - cost) * ((R)1.0 + cost));
  const R phi = (R)k2Pi * (R)rng.flat();
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  secDir[0] = sint * cosPhi;
  secDir[1] = sint * sinPhi;
  secDir[2] = cost;
  Rotate(secDir, primDir);
  const R me = (R)kElectronMassC2;
  const R primETot = primEkin + me;
  const R primPTot = dsqrt(primEkin * (primETot + me));
  primDir[0] = primPTot * primDir[0] - secEkin * secDir[0];
  primDir[1] = primPTot * primDir[1] - secEkin * secDir[1];
  primDir[2] = primPTot * primDir[2] - secEkin * secDir[2];
  const R norm = (R)1.0 / dsqrt(primDir[0] * primDir[0] + primDir[1] * primDir[1] + primDir[2] * primDir[2]);
  primDir[0] *= norm;
  primDir[1] *= norm;
  primDir[2] *= norm;
}
template <typename R> HD void BremPerform(const DevTables<R>& T, Trk<R>& t, SecList<R>& sl, Rng& rng, bool isSB) {  // :35-70
  const double primEkin = t.ekin;
  const R lEkin = t.logE();
  const R pE = (R)primEkin;
  if (pE <= (R)T.mcGamCut[t.imc]) return;
  const R eGamma = isSB ? SampleETransferSB(T, primEkin, lEkin, t.imc, rng, t.isEl)
                        : SampleETransferRB(T, primEkin, lEkin, t.imc, rng, t.isEl);
  Sec<R>* s = &sl.s0;
  sl.ns = 1;
  s->ptype = 2;
  BremSampleDirections(pE, eGamma, s->dir, t.dir, rng);
  t.setE(primEkin - (double)eGamma);
  s->ekin = (double)eGamma;
}

// ---- G4HepEmPositronInteractionAnnihilation.icc ------------------------------------------------
template <typename R> HD void AnnihilateAtRest(SecList<R>& sl, Rng& rng) {  // :23-49
  const R cost = (R)2. * (R)rng.flat() - (R)1.;
  const R sint = dsqrt(((R)1. - cost) * ((R)1. + cost));
  const R phi = (R)k2Pi * (R)rng.flat();
  Sec<R>* g1 = &sl.s0;
  Sec<R>* g2 = &sl.s1;
  sl.ns = 2;
  g1->ptype = g2->ptype = 2;
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  g1->dir[0] = sint * cosPhi;
  g1->dir[1] = sint * sinPhi;
  g1->dir[2] = cost;
  g2->dir[0] = -g1->dir[0];
  g2->dir[1] = -g1->dir[1];
  g2->dir[2] = -g1->dir[2];
  g1->ekin = kElectronMassC2;
  g2->ekin = kElectronMassC2;
}
template <typename R> HD void AnnihilateInFlight(Trk<R>& t, SecList<R>& sl, Rng& rng) {  // :51-119
  const R primEkin = (R)t.ekin;
  const R* primDir = t.dir;
  Sec<R>* g1 = &sl.s0;
  Sec<R>* g2 = &sl.s1;
  sl.ns = 2;
  g1->ptype = g2->ptype = 2;
  const R me = (R)kElectronMassC2;
  const R tau = primEkin * (R)kInvElectronMassC2;
  const R gam = tau + (R)1.0;
  const R tau2 = tau + (R)2.0;
  const R sqgrate = dsqrt(tau / tau2) * (R)0.5;
  const R epsmin = (R)0.5 - sqgrate, epsmax = (R)0.5 + sqgrate;
  const R epsqot = epsmax / epsmin;
  const R tau4 = tau2 * tau2;
  R eps, rfunc, r[2];
  do {
    rng.flatArray(2, r);
    eps = epsmin * vexp(vlog(epsqot) * r[0]);
    rfunc = (R)1. - eps + ((R)2. * gam * eps - (R)1.) / (eps * tau4);
  } while (rfunc < r[1]);
  const R sqg2m1 = dsqrt(tau * tau2);
  const R cost = tmin((R)1., tmax((R)-1., (eps * tau2 - (R)1.) / (eps * sqg2m1)));
  const R sint = dsqrt(((R)1. + cost) * ((R)1. - cost));
  const R phi = (R)k2Pi * (R)rng.flat();
  const R initEt = primEkin + (R)2. * me;
  const R ekinG1 = eps * initEt;
  g1->ekin = (double)ekinG1;
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  g1->dir[0] = sint * cosPhi;
  g1->dir[1] = sint * sinPhi;
  g1->dir[2] = cost;
  Rotate(g1->dir, primDir);
  g2->ekin = (double)(initEt - ekinG1);
  const R initPt = dsqrt(primEkin * (primEkin + (R)2 * me));
  const R px = initPt * primDir[0] - g1->dir[0] * ekinG1;
  const R py = initPt * primDir[1] - g1->dir[1] * ekinG1;
  const R pz = initPt * primDir[2] - g1->dir[2] * ekinG1;
  const R norm = (R)1.0 / dsqrt(px * px + py * py + pz * pz);
  g2->dir[0] = px * norm;
  g2->dir[1] = py * norm;
  g2->dir[2] = pz * norm;
  t.setE(0.0);
}

// ---- G4HepEmElectronManager.icc: stepping ----------------------------------------------------
template <typename R> struct ElState {
  Trk<R> t;
  Msc<R> msc;
  R range, pstep;
  double preEkin;
  R preLEkin;
};

template <typename R>
HD void ConvertTrueToGeometricLength(const DevTables<R>& T, Msc<R>& m, R ekin, R lekin, R range, int imc, bool isEl) {  // :602-650
  m.par1 = (R)-1.;
  m.par2 = (R)0.;
  m.par3 = (R)0.;
  m.trueStep = tmin(m.trueStep, range);
  m.zpath = m.trueStep;
  if (m.trueStep < (R)1.0E-6) return;
  const R tau = m.trueStep / m.lambtr1;
  if (tau < (R)1.0e-16) {
    m.zpath = tmin(m.trueStep, m.lambtr1);
  } else if (m.trueStep < range * (R)0.05) {
    m.zpath = (tau < (R)1.0e-6) ? m.trueStep * ((R)1. - (R)0.5 * tau) : m.lambtr1 * ((R)1. - vexp(-tau));
  } else if (ekin < (R)kElectronMassC2 || m.trueStep == range) {
    m.par1 = (R)1. / range;
    m.par2 = (R)1. / (m.par1 * m.lambtr1);
    m.par3 = (R)1. + m.par2;
    m.zpath = (R)1. / (m.par1 * m.par3);
    if (m.trueStep < range) m.zpath *= ((R)1. - vpow((R)1. - m.trueStep / range, m.par3));
  } else {
    const R rfin = tmax(range - m.trueStep, (R)0.01 * range);
    const DevEl<R>& e = T.el[isEl ? 0 : 1];
    const double t1 = GetInvRange(e, imc, (double)rfin, LogBin(lekin, e.lossLogMin, e.lossILD, e.N));
    const int imat = T.mcMat[imc];
    const R lambda1 = GetTransportMFP(e, imat, t1, vlog((R)t1));
    m.par1 = (m.lambtr1 - lambda1) / (m.lambtr1 * m.trueStep);
    m.par2 = (R)1. / (m.par1 * m.lambtr1);
    m.par3 = (R)1. + m.par2;
    m.zpath = ((R)1. - vpow(lambda1 / m.lambtr1, m.par3)) / (m.par1 * m.par3);
  }
  m.zpath = tmin(m.zpath, m.lambtr1);
}

template <typename R> HD void ConvertGeometricToTrueLength(Msc<R>& m, R range, R gStep) {  // :653-690
  m.zpath = gStep;
  if (gStep < (R)1.0E-6) {
    m.trueStep = gStep;
    return;
  }
  R tlength = gStep;
  if (gStep > m.lambtr1 * (R)1.0e-16) {
    if (m.par1 < (R)0.) {
      tlength = -m.lambtr1 * vlog((R)1. - gStep / m.lambtr1);
    } else {
      const R dum = m.par1 * m.par3 * gStep;
      tlength = (dum < (R)1.) ? ((R)1. - vpow((R)1. - dum, (R)1. / m.par3)) / m.par1 : range;
    }
    if (tlength < gStep) tlength = gStep;
  }
  m.trueStep = tlength;
}

template <typename R> HD void HowFarToDiscreteInteraction(const DevTables<R>& T, ElState<R>& el) {  // :48-101
  Trk<R>& t = el.t;
  int winner = -1;
  const double ekin = t.ekin;
  const R lekin = t.logE();
  const int imc = t.imc;
  const DevEl<R>& e = T.el[t.isEl ? 0 : 1];
  const R range = GetRestRange(e, imc, ekin, lekin);
  el.range = range;
  const int ireg = T.mcReg[imc];
  const R frange = (R)T.regFinalRange[ireg], drange = (R)T.regDRover[ireg];
  R pStep = (range > frange) ? range * drange + frange * ((R)1.0 - drange) * ((R)2.0 - frange / range) : range;
  const int imat = T.mcMat[imc];
  R mx[4];
  mx[0] = GetRestMacXSecForStepping(e, imc, ekin, lekin, true);
  mx[1] = GetRestMacXSecForStepping(e, imc, ekin, lekin, false);
  mx[2] = t.isEl ? (R)0.0 : MacXsecAnnihilation((R)(0.8 * ekin), (R)T.matElDens[imat]);
  mx[3] = GetMacXSecNuclear(e, imat, ekin, lekin);
#pragma unroll
  for (int ip = 0; ip < 4; ++ip) {
    const R mfp = (mx[ip] > (R)0.) ? (R)1. / mx[ip] : (R)kALargeValue;
    t.mfp[ip] = mfp;
    const R dStep = mfp * t.nial[ip];
    if (dStep < pStep) {
      pStep = dStep;
      winner = ip;
    }
  }
  el.pstep = pStep;
  t.winner = winner;
  t.gstep = pStep;
}

template <typename R> HD void HowFarToMSC(const DevTables<R>& T, ElState<R>& el, Rng& rng) {  // :103-164
  R pStep = el.pstep;
  const R range = el.range;
  Trk<R>& t = el.t;
  const R ekin = (R)t.ekin;
  const R lekin = t.logE();
  const int imc = t.imc;
  const DevEl<R>& e = T.el[t.isEl ? 0 : 1];
  const int imat = T.mcMat[imc], ireg = T.mcReg[imc];
  const double ekinD = t.ekin;
  Msc<R>& m = el.msc;
  m.trueStep = pStep;
  m.zpath = pStep;
  m.active = false;
  m.displ[0] = m.displ[1] = m.displ[2] = (R)0.;
  m.dir[0] = (R)0.;
  m.dir[1] = (R)0.;
  m.dir[2] = (R)1.;
  if (pStep > (R)5.E-8 && ekin > (R)1.0E-3) {
    m.active = true;
    m.lambtr1 = GetTransportMFP(e, imat, ekinD, lekin);
    UMSCStepLimit(T, m, ekin, imat, ireg, range, t.safety, t.onBoundary, t.isEl, rng);
    ConvertTrueToGeometricLength(T, m, ekin, lekin, range, imc, t.isEl);
    const R mscTrue = m.trueStep;
    if (mscTrue < pStep) {
      t.winner = -2;
      pStep = mscTrue;
      el.pstep = pStep;
    }
    t.gstep = tmin(m.zpath, pStep);
  }
}

template <typename R> HD void UpdatePStepLength(ElState<R>& el) {  // :171-203
  const R gStep = el.t.gstep;
  R pStep = gStep;
  const R range = el.range;
  Msc<R>& m = el.msc;
  if (m.active) {
    pStep = m.trueStep;
    if (gStep < m.zpath) {
      ConvertGeometricToTrueLength(m, range, gStep);
      pStep = tmin(pStep, m.trueStep);
      m.trueStep = pStep;
    }
    if (pStep <= (R)5.E-8 || range <= pStep) m.active = false;
  }
  el.pstep = pStep;
}

template <typename R> HD void UpdateNumIALeft(ElState<R>& el) {  // :205-214
  const R pStep = el.pstep;
#pragma unroll
  for (int k = 0; k < 4; ++k) el.t.nial[k] -= pStep / el.t.mfp[k];
}

// :216-259; the mean-loss update near the end of the range is an FP64 island
template <typename R> HD bool ApplyMeanEnergyLoss(const DevTables<R>& T, ElState<R>& el) {
  const R pStep = el.pstep;
  Trk<R>& t = el.t;
  const double ekin = t.ekin;
  const R range = el.range;
  // one store site per field on every path (stores of different fields on
  // the two sides of a branch get merged into an indexed store, which puts
  // the track state in scratch memory)
  double eloss = ekin;
  bool stopped = true;
  if (!(pStep >= range || ekin <= T.minLossE)) {
    const DevEl<R>& e = T.el[t.isEl ? 0 : 1];
    const int imc = t.imc;
    const R lekin = t.logE();
    eloss = (double)(pStep * GetRestDEDX(e, imc, ekin, lekin));
    const double linLim = T.regLinLim[T.mcReg[imc]];
    if (eloss > ekin * linLim) {
      const double postRange = (double)range - (double)pStep;
      eloss = ekin - GetInvRange(e, imc, postRange, LogBin(lekin, e.lossLogMin, e.lossILD, e.N));
    }
    eloss = tmax(eloss, 0.0);
    stopped = eloss >= ekin;
    if (stopped) eloss = ekin;
  }
  t.setE(stopped ? 0.0 : ekin - eloss);
  t.edep = eloss;
  return stopped;
}

template <typename R> HD void SampleMSC(const DevTables<R>& T, ElState<R>& el, Rng& rng) {  // :261-322
  const R pStep = el.pstep;
  Trk<R>& t = el.t;
  const int imc = t.imc;
  const R preEkin = (R)el.preEkin;
  const R range = el.range;
  Msc<R>& m = el.msc;
  if (m.active && (pStep > tmax((R)1.0E-8, (R)1.0e-16 * m.lambtr1))) {
    double postEkinD = el.preEkin;
    R postLEkin = el.preLEkin;
    if (pStep > range * (R)0.01) {
      postEkinD = t.ekin;
      postLEkin = t.logE();
    }
    const R postEkin = (R)postEkinD;
    const int imat = T.mcMat[imc];
    const R postTr1mfp = GetTransportMFP(T.el[t.isEl ? 0 : 1], imat, postEkinD, postLEkin);
    UMSCSampleScattering(T, m, pStep, preEkin, m.lambtr1, postEkin, postTr1mfp, imat, t.isEl, T.posCor != 0, T.displace != 0, rng);
    if (!m.noScat) {
      Rotate(m.dir, t.dir);
      if (m.isDisplace) Rotate(m.displ, t.dir);
      t.dir[0] = m.dir[0];
      t.dir[1] = m.dir[1];
      t.dir[2] = m.dir[2];
    }
  }
}

template <typename R> HD bool SampleLossFluctuations(const DevTables<R>& T, ElState<R>& el, Rng& rng) {  // :324-368
  Trk<R>& t = el.t;
  const int imc = t.imc;
  const double preEkin = el.preEkin;
  double finalEkin = t.ekin;
  double eloss = t.edep;
  const int ireg = T.mcReg[imc];
  if (T.regFluct[ireg] && eloss > 1.E-5) {
    const double elCut = T.mcElCut[imc];
    const double meanExE = T.matMeanExc[T.mcMat[imc]];
    const double tmx = t.isEl ? 0.5 * preEkin : preEkin;
    const double tcut = tmin(elCut, tmx);
    // in the FP32 form the collision numbers are float: the remainder
    // a3 - namean then carries an absolute error of ~6e-8 a3 (< 1e-2 of a
    // Poisson mean of ~8 up to a3 ~ 1e5); the energies around it stay double
    eloss = (double)SampleELossFluctuation((R)tcut, (R)meanExE, (R)eloss, rng);
    eloss = tmax(eloss, 0.0);
    finalEkin = preEkin - eloss;
  }
  const bool stopped = finalEkin <= T.elTrackCut;
  if (stopped) {
    eloss = preEkin;
    finalEkin = 0.0;
  }
  t.setE(finalEkin);
  t.edep = eloss;
  return stopped;
}

template <typename R> HD bool CheckDelta(const DevTables<R>& T, Trk<R>& t, R rnd) {  // :408-423
  const DevEl<R>& e = T.el[t.isEl ? 0 : 1];
  const int iDProc = t.winner;
  const int imc = t.imc;
  const int imat = T.mcMat[imc];
  const double ekin = t.ekin;
  const R lekin = t.logE();
  const R mxsec = iDProc < 2 ? GetRestMacXSec(e, imc, ekin, lekin, iDProc == 0)
                : (iDProc < 3 ? MacXsecAnnihilation((R)ekin, (R)T.matElDens[imat]) : GetMacXSecNuclear(e, imat, ekin, lekin));
  return mxsec <= (R)0.0 || rnd > mxsec * PickMFP(t.mfp, iDProc);
}

// PerformDiscrete (:425-441) up to the final-state sampling: the
// interaction-length reset and the delta-interaction test. Returns the
// process whose final state is to be sampled (0 ioni, 1 brem, 2 annihilation,
// 3 nuclear: no final state) or -1.
template <typename R> HD int ElectronSelectDiscrete(const DevTables<R>& T, Trk<R>& t, Rng& rng) {
  const int iDProc = t.winner;
  if (iDProc < 0 || t.onBoundary) return -1;
  ResetNumIALeft(t.nial, iDProc);
  if (CheckDelta(T, t, (R)rng.flat())) return -1;
  return iDProc;
}

// G4HepEmElectronManager::HowFar(data, pars, tlData) (:35-45, 166-169)
template <typename R> HD void ElectronHowFar(const DevTables<R>& T, ElState<R>& el, Rng& rng) {
#pragma unroll
  for (int ip = 0; ip < 4; ++ip)
    if (el.t.nial[ip] <= (R)0.) el.t.nial[ip] = -vlog((R)rng.flat());
  HowFarToDiscreteInteraction(T, el);
  HowFarToMSC(T, el, rng);
}

// G4HepEmElectronManager::Perform(data, pars, tlData) (:461-483) up to
// PerformDiscrete: PerformContinuous (:375-405). Returns 1 if the track
// stopped in the step, 0 if it goes on to PerformDiscrete and -1 for a zero
// geometrical step (nothing else happens in the step).
template <typename R> HD int ElectronPerformContinuous(const DevTables<R>& T, ElState<R>& el, Rng& rng) {
  Trk<R>& t = el.t;
  t.edep = 0;
  el.pstep = t.gstep;
  if (t.gstep <= (R)0.) return -1;
  el.preEkin = t.ekin;
  el.preLEkin = t.logE();
  UpdatePStepLength(el);
  bool stopped = false;
  if (el.pstep > (R)0.0) {  // PerformContinuous returns "not stopped" on a zero step
    UpdateNumIALeft(el);
    stopped = ApplyMeanEnergyLoss(T, el);
    if (!stopped) {
      SampleMSC(T, el, rng);
      stopped = SampleLossFluctuations(T, el, rng);
    }
  }
  return stopped ? 1 : 0;
}

// ---- G4HepEmGammaManager.icc ------------------------------------------------------------------
template <typename R> HD R SandiaEval(const double* energies, const double* coefs, int nint, R ekin) {
  int interval = 0;
  if (ekin >= (R)energies[0]) {
    for (int i = nint - 1; i >= 0; i--) {
      if (ekin >= (R)energies[i]) { interval = i; break; }
    }
  }
  const double* c = coefs + 4 * interval;
  const R inv = (R)1 / ekin;
  return inv * ((R)c[0] + inv * ((R)c[1] + inv * ((R)c[2] + inv * (R)c[3])));
}
template <typename R> HD R GetMacXSecPE(const DevTables<R>& T, int imat, R ekin) {  // :155-170
  const int s = T.matSandiaStart[imat];
  return SandiaEval(T.matSandiaE + s, T.matSandiaC + 4 * s, T.matSandiaN[imat], ekin);
}
template <typename R> HD R GetTotalMacXSec(const DevTables<R>& T, Trk<R>& t, R& peMxSec) {  // :108-153
  const double ekin = t.ekin;
  const R lekin = t.logE();
  const int imat = T.mcMat[t.imc];
  if (ekin > T.gEMax1) {
    const int off = imat * T.gmPerMat + T.gmOff2;
    return SplineLog4(T.gmN2, T.gmData64 + off, T.gmData + off, ekin, lekin, T.gLogEMin2, T.gILD2, 1);
  }
  if (ekin > T.gEMax0) {
    const int off = imat * T.gmPerMat + T.gmOff1;
    R mx[2];
    LinearLog2All(T.gmN1, T.gmData64 + off, T.gmData + off, ekin, lekin, T.gLogEMin1, T.gILD1, mx);
    peMxSec = mx[1];
    return mx[0];
  }
  const int off = imat * T.gmPerMat;
  const R comp = LinearLog(T.gmN0, T.gmData64 + off, T.gmData + off, ekin, lekin, T.gLogEMin0, T.gILD0);
  const R pe = tmax((R)0.0, GetMacXSecPE(T, imat, (R)ekin));
  peMxSec = pe;
  return comp + pe;
}
// :187-219. In the gamma-nuclear bin the reference reads a fifth value past the
// stored four; it only feeds an unused PE cross section, so it is not read here.
template <typename R> HD void SampleInteraction(const DevTables<R>& T, Trk<R>& t, R& peMxSec, R urnd) {
  const R tmfp = t.mfp[0];
  t.nial[0] = (R)-1.0;
  if (t.ekin > T.gEMax1) {
    const double ekin = t.ekin;
    const R lekin = t.logE();
    const int off = T.mcMat[t.imc] * T.gmPerMat + T.gmOff2;
    R mxSec = (R)0.0, cProb = (R)0.0;
    int pid = 2;
    do {
      if (pid == 5) { ++pid; break; }
      mxSec = SplineLog4(T.gmN2, T.gmData64 + off, T.gmData + off, ekin, lekin, T.gLogEMin2, T.gILD2, pid);
      cProb += mxSec * tmfp;
      ++pid;
    } while (pid < 6 && urnd > cProb);
    t.winner = pid - 3;
    peMxSec = mxSec;
    return;
  }
  t.winner = (urnd > tmfp * peMxSec) ? 1 : 2;
}

// ---- gamma interactions -------------------------------------------------------------------
template <typename R> HD void ComptonPerform(Trk<R>& t, SecList<R>& sl, Rng& rng) {  // Compton.icc:17-103
  const double primEd = t.ekin;
  const R primE = (R)primEd;
  const R thr = (R)0.0001;
  if (primE < thr) return;
  const R org[3] = {t.dir[0], t.dir[1], t.dir[2]};
  const R kappa = primE * (R)kInvElectronMassC2;
  const R eps0 = (R)1. / ((R)1. + (R)2. * kappa);
  const R eps02 = eps0 * eps0;
  const R al1 = -vlog(eps0);
  const R al2 = al1 + (R)0.5 * ((R)1. - eps02);
  R eps, eps2, gf, oneMinusCost, sint2, r[3];
  do {
    rng.flatArray(3, r);
    if (al1 > al2 * r[0]) {
      eps = vexp(-al1 * r[1]);
      eps2 = eps * eps;
    } else {
      eps2 = eps02 + ((R)1. - eps02) * r[1];
      eps = dsqrt(eps2);
    }
    oneMinusCost = ((R)1. - eps) / (eps * kappa);
    sint2 = oneMinusCost * ((R)2. - oneMinusCost);
    gf = (R)1. - eps * sint2 / ((R)1. + eps2);
  } while (gf < r[2]);
  const R cost = (R)1.0 - oneMinusCost;
  const R sint = dsqrt(tmax((R)0., sint2));
  const R phi = (R)k2Pi * (R)rng.flat();
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  t.dir[0] = sint * cosPhi;
  t.dir[1] = sint * sinPhi;
  t.dir[2] = cost;
  Rotate(t.dir, org);
  const R postE = primE * eps;
  const R secE = primE - postE;
  double edep = 0.0;
  if (secE > thr) {
    Sec<R>* s = &sl.s0;
    sl.ns = 1;
    s->ptype = 0;
    s->dir[0] = primE * org[0] - postE * t.dir[0];
    s->dir[1] = primE * org[1] - postE * t.dir[1];
    s->dir[2] = primE * org[2] - postE * t.dir[2];
    const R norm = (R)1.0 / dsqrt(s->dir[0] * s->dir[0] + s->dir[1] * s->dir[1] + s->dir[2] * s->dir[2]);
    s->dir[0] *= norm;
    s->dir[1] *= norm;
    s->dir[2] *= norm;
    s->ekin = (double)secE;
  } else {
    edep += (double)secE;
  }
  if (postE > thr) {
    t.setE((double)postE);
  } else {
    edep += (double)postE;
    t.setE(0.0);
  }
  t.edep = edep;
}

template <typename R>
HD int ConvSelectTargetAtom(const DevTables<R>& T, int imat, double ekin, R lekin, R urndn) {  // Conversion.icc:151-176
  const R* data = T.convSelData + T.convSelStart[imat];
  const int numData = T.convSelN;
  const int numElem = (int)data[0];
  const double* xd = T.convSelGrid64;
  const double xv = tmax(xd[0], tmin(xd[numData - 1], ekin));
  const int idx = (int)tmax((R)0.0, tmin((lekin - T.convSelLogMin) * T.convSelILD, (R)(numData - 2)));
  const double x1 = xd[idx], x2 = xd[idx + 1];
  const double dl = x2 - x1;
  const R b = Frac(xv - x1, dl, R());
  const int i0 = idx * (numElem - 1) + 1;
  const int i1 = i0 + (numElem - 1);
  int ie = 0;
  while (ie < numElem - 1 && urndn > data[i0 + ie] + b * (data[i1 + ie] - data[i0 + ie])) ++ie;
  return ie;
}
template <typename R> HD void ConvScreen12(R delta, R& f1, R& f2) {  // Conversion.icc:266-274
  if (delta > (R)1.4) {
    f1 = (R)42.038 - (R)8.29 * vlog(delta + (R)0.958);
    f2 = f1;
  } else {
    f1 = (R)42.184 - delta * ((R)7.444 - (R)1.623 * delta);
    f2 = (R)41.326 - delta * ((R)5.848 - (R)0.902 * delta);
  }
}
template <typename R> HD void ConvPhi12(R delta, R& phi1, R& phi2) {  // Conversion.icc:240-248
  if (delta > (R)1.4) {
    phi1 = (R)21.0190 - (R)4.145 * vlog(delta + (R)0.958);
    phi2 = phi1;
  } else {
    phi1 = (R)20.806 - delta * ((R)3.190 - (R)0.5710 * delta);
    phi2 = (R)20.234 - delta * ((R)2.126 - (R)0.0903 * delta);
  }
}
HD double cbrt13(double x) { return pow(x, 1. / 3.); }
HD float cbrt13(float x) { return cbrtf(x); }
template <typename R> HD void ConvPerform(const DevTables<R>& T, Trk<R>& t, SecList<R>& sl, Rng& rng) {  // Conversion.icc:19-146
  const double primEd = t.ekin;
  const R primE = (R)primEd;
  if (primEd < 2. * kElectronMassC2) return;
  const R lPrimE = t.logE();
  const int imat = T.mcMat[t.imc];
  const int nel = T.matNElem[imat];
  const int elemIndx = nel > 1 ? ConvSelectTargetAtom(T, imat, primEd, lPrimE, (R)rng.flat()) : 0;
  const int iZet = T.matElemZ[T.matElemStart[imat] + elemIndx];
  const R lpmEnr = (R)kLPMconstant * (R)T.matRadLen[imat];
  const int z = iZet < T.maxZ ? iZet : T.maxZ;
  const R me = (R)kElectronMassC2;
  const R eps0 = me / primE;
  R eps;
  if (primE < (R)2.0) {
    eps = eps0 + ((R)0.5 - eps0) * (R)rng.flat();
  } else {
    const R deltaFactor = eps0 * (R)136. / (R)T.elZ13[z];
    const R deltaMin = (R)4. * deltaFactor;
    const R deltaMax = (primE < (R)50.0) ? (R)T.elDMaxLow[z] : (R)T.elDMaxHigh[z];
    const R logZ13 = (R)0.333333 * (R)T.elLogZ[z];
    const R FZ = (primE < (R)50.0) ? (R)8. * logZ13 : (R)8. * (logZ13 + (R)T.elCoulomb[z]);
    const R epsp = (R)0.5 - (R)0.5 * dsqrt((R)1. - deltaMin / deltaMax);
    const R epsMin = tmax(eps0, epsp);
    const R epsRange = (R)0.5 - epsMin;
    R F10, F20;
    ConvScreen12(deltaMin, F10, F20);
    F10 -= FZ;
    F20 -= FZ;
    const R NormF1 = tmax(F10 * epsRange * epsRange, (R)0.);
    const R NormF2 = tmax((R)1.5 * F20, (R)0.);
    const R NormCond = NormF1 / (NormF1 + NormF2);
    const R invF10 = (R)1. / F10, invF20 = (R)1. / F20;
    R r[3], greject;
    if (primE < (R)100000.0) {  // SampleEnergyRateNoLPM :179-200
      // the two branches share delta and, above delta = 1.4, the screening
      // function: written once, so a wave with lanes in both branches
      // evaluates them once (same operations per lane as the branches)
      do {
        rng.flatArray(3, r);
        const bool first = NormCond > r[0];
        if (first) eps = (R)0.5 - epsRange * cbrt13(r[1]);
        else eps = epsMin + epsRange * r[1];
        const R delta = deltaFactor / (eps * ((R)1. - eps));
        const R screen = delta > (R)1.4 ? (R)42.038 - (R)8.29 * vlog(delta + (R)0.958)
                         : first        ? (R)42.184 - delta * ((R)7.444 - (R)1.623 * delta)
                                        : (R)41.326 - delta * ((R)5.848 - (R)0.902 * delta);
        greject = (screen - FZ) * (first ? invF10 : invF20);
      } while (greject < r[2]);
    } else {  // SampleEnergyRateWithLPM :203-237
      const R z23 = (R)T.elZ23[z], ilVarS1 = (R)T.elILS1[z], ilVarS1Cond = (R)T.elILS1Cond[z];
      do {
        rng.flatArray(3, r);
        R fXi, fG, fPhi, phi1, phi2;
        if (NormCond > r[0]) {
          eps = (R)0.5 - epsRange * cbrt13(r[1]);
          const R delta = deltaFactor / (eps * ((R)1. - eps));
          ConvPhi12(delta, phi1, phi2);
          EvaluateLPMFunctions(fXi, fG, fPhi, primE, eps * primE, lpmEnr, z23, ilVarS1, ilVarS1Cond, (R)0.0, (R)-1.0);
          greject = fXi * (((R)2. * fPhi + fG) * phi1 - fG * phi2 - fPhi * FZ) * invF10;
        } else {
          eps = epsMin + epsRange * r[1];
          const R delta = deltaFactor / (eps * ((R)1. - eps));
          ConvPhi12(delta, phi1, phi2);
          EvaluateLPMFunctions(fXi, fG, fPhi, primE, eps * primE, lpmEnr, z23, ilVarS1, ilVarS1Cond, (R)0.0, (R)-1.0);
          greject = fXi * ((fPhi + (R)0.5 * fG) * phi1 + (R)0.5 * fG * phi2 - (R)0.5 * (fG + fPhi) * FZ) * invF20;
        }
      } while (greject < r[2]);
    }
  }
  R eTot, pTot;
  if ((R)rng.flat() > (R)0.5) {
    eTot = ((R)1. - eps) * primE;
    pTot = eps * primE;
  } else {
    pTot = ((R)1. - eps) * primE;
    eTot = eps * primE;
  }
  const R eKin = tmax((R)0., eTot - me);
  const R pKin = tmax((R)0., pTot - me);
  Sec<R>* se = &sl.s0;
  Sec<R>* sp = &sl.s1;
  sl.ns = 2;
  se->ptype = 0;
  sp->ptype = 1;
  const R phi = (R)k2Pi * (R)rng.flat();
  R sinPhi, cosPhi;
  dsincos(phi, sinPhi, cosPhi);
  const R costEl = SampleCostModifiedTsai(eKin, rng);
  const R sintEl = dsqrt(((R)1.0 - costEl) * ((R)1.0 + costEl));
  se->dir[0] = sintEl * cosPhi;
  se->dir[1] = sintEl * sinPhi;
  se->dir[2] = costEl;
  Rotate(se->dir, t.dir);
  const R costPos = SampleCostModifiedTsai(pKin, rng);
  const R sintPos = dsqrt(((R)1.0 - costPos) * ((R)1.0 + costPos));
  sp->dir[0] = -sintPos * cosPhi;
  sp->dir[1] = -sintPos * sinPhi;
  sp->dir[2] = costPos;
  Rotate(sp->dir, t.dir);
  se->ekin = (double)eKin;
  sp->ekin = (double)pKin;
  t.setE(0.0);
}

template <typename R> HD void PEPerform(const DevTables<R>& T, Trk<R>& t, R peMxSec, SecList<R>& sl, Rng& rng) {  // Photoelectric.icc:13-121
  const double gEd = t.ekin;
  const R gE = (R)gEd;
  // SelectElementBindingEnergy :39-78
  const int imat = T.mcMat[t.imc];
  const int nel = T.matNElem[imat];
  const int est = T.matElemStart[imat];
  int ielem = 0;
  if (nel > 1) {
    const R x = (R)rng.flat() * peMxSec;
    R sum = (R)0;
    const R invE = (R)1 / gE;
    for (int i = 0; i < nel; i++) {
      const int Z = T.matElemZ[est + i];
      const int s = T.elSandiaStart[Z];
      const double* en = T.elSandiaE + s;
      int interval = 0;
      if (gE >= (R)en[0]) {
        for (int j = T.elSandiaN[Z] - 1; j >= 0; j--) {
          if (gE >= (R)en[j]) { interval = j; break; }
        }
      }
      const double* c = T.elSandiaC + 4 * (s + interval);
      sum += (R)T.matElemNat[est + i] * invE * ((R)c[0] + invE * ((R)c[1] + invE * ((R)c[2] + invE * (R)c[3])));
      if (x <= sum) { ielem = i; break; }
    }
  }
  const double be = T.elKBind[T.matElemZ[est + ielem]];
  const double peE = gEd - be;
  if (peE > 0.000001) {
    Sec<R>* s = &sl.s0;
    sl.ns = 1;
    s->ptype = 0;
    const R kinE = (R)peE;
    const R tau = kinE * (R)kInvElectronMassC2;
    const R gamma = (R)1.0 + tau;
    const R beta = dsqrt(tau * (tau + (R)2.0)) / gamma;
    const R ac = ((R)1.0 - beta) / beta;
    const R a1 = (R)0.5 * beta * gamma * tau * (gamma - (R)2.0);
    const R a2 = ac + (R)2.0;
    const R gtmax = (R)2.0 * (a1 + (R)1.0 / ac);
    R tsam, gtr, r[2];
    do {
      rng.flatArray(2, r);
      tsam = (R)2.0 * ac * ((R)2.0 * r[0] + a2 * dsqrt(r[0])) / (a2 * a2 - (R)4.0 * r[0]);
      gtr = ((R)2.0 - tsam) * (a1 + (R)1.0 / (ac + tsam));
    } while (r[1] * gtmax > gtr);
    const R costheta = (R)1.0 - tsam;
    const R sint = dsqrt(tsam * ((R)2.0 - tsam));
    const R phi = (R)k2Pi * (R)rng.flat();
    R sinPhi, cosPhi;
    dsincos(phi, sinPhi, cosPhi);
    s->dir[0] = sint * cosPhi;
    s->dir[1] = sint * sinPhi;
    s->dir[2] = costheta;
    Rotate(s->dir, t.dir);
    s->ekin = peE;
    t.edep = be;
  } else {
    t.edep = gEd;
  }
  t.setE(0.0);
}

}  // namespace hepem
