// hepem_tables.cpp -- offline builder of G4HepEmData / G4HepEmParameters in
// the reference layout, from analytic physics models (no Geant4).
//
// Array shapes, grids and index conventions follow the reference builders:
//   loss grid + [R,R''],[dEdx,dEdx''],[InvR''] per couple
//       G4HepEm/G4HepEmInit/src/G4HepEmElectronTableBuilder.cc:40-178
//   restricted macroscopic cross sections [M,Emax,max,lnE0,1/dlnE,(E,S,S'')xM]
//       G4HepEmElectronTableBuilder.cc:184-330
//   element selectors [K,Q,lnE0,1/dlnE,(E,P_0..P_{Q-2})xK]
//       G4HepEmElectronTableBuilder.cc:501-682
//   Seltzer-Berger sampling tables per Z [#data,minE,maxE,#gcuts,(cuts,(cum,A,B)x54) per E]
//       G4HepEmElectronTableBuilder.cc:685-834, G4HepEmSBBremTableBuilder.cc:60-300
//   gamma windows 32/32/256 and the conversion element selector
//       G4HepEm/G4HepEmInit/src/G4HepEmGammaTableBuilder.cc:30-286
//   element / material / couple data
//       G4HepEm/G4HepEmInit/src/G4HepEmMaterialInit.cc:30-255
// The physics inside (cross sections, stopping powers, sampling tables) is
// this file's own: textbook formulas cited inline. All arithmetic is plain
// double with no contraction, so the tables are bit-reproducible.

#include "g4hepem_amd_tables.h"
#include "g4hepem_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// ---- constants (Geant4/CLHEP internal units: MeV, mm) ----------------------
constexpr double kPi     = 3.14159265358979323846;
constexpr double kMe     = 0.51099891;               // electron mass [MeV]
constexpr double kRe     = 2.8179402894e-12;         // classical electron radius [mm]
constexpr double kAlpha  = 7.2973525653e-3;          // fine structure constant
constexpr double kHbarc  = 1.973269631e-10;          // hbar*c [MeV mm]
constexpr double kBohr   = 5.2917720859e-8;          // Bohr radius [mm]
constexpr double kNA     = 6.02214179e23;            // Avogadro [1/mol]
constexpr double kEV     = 1.0e-6;                   // 1 eV in MeV
constexpr double kKeV    = 1.0e-3;
constexpr double kBarn   = 1.0e-22;                  // 1 barn in mm^2
constexpr double kMigdal = 4.0 * kPi * kRe * (kHbarc / kMe) * (kHbarc / kMe);

// ---- small material database -----------------------------------------------
struct MatDef {
  const char* name;
  double rho;      // g/cm3
  double meanExc;  // eV
  int    nel;
  int    Z[4];
  double count[4]; // atoms per formula unit
};

const MatDef kMaterials[] = {
  {"G4_Pb",     11.35,  823.0, 1, {82, 0, 0, 0}, {1, 0, 0, 0}},
  {"G4_lAr",    1.396,  188.0, 1, {18, 0, 0, 0}, {1, 0, 0, 0}},
  {"G4_Si",     2.33,   173.0, 1, {14, 0, 0, 0}, {1, 0, 0, 0}},
  {"G4_PbWO4",  8.28,   600.7, 3, {82, 74, 8, 0}, {1, 1, 4, 0}},
  {"G4_W",      19.3,   727.0, 1, {74, 0, 0, 0}, {1, 0, 0, 0}},
  {"G4_Cu",     8.96,   322.0, 1, {29, 0, 0, 0}, {1, 0, 0, 0}},
  {"G4_Fe",     7.874,  286.0, 1, {26, 0, 0, 0}, {1, 0, 0, 0}},
  {"G4_WATER",  1.0,     78.0, 2, {1, 8, 0, 0},  {2, 1, 0, 0}},
};

double AtomicMass(int Z) {  // g/mol
  switch (Z) {
    case 1:  return 1.00794;
    case 8:  return 15.9994;
    case 14: return 28.0855;
    case 18: return 39.948;
    case 26: return 55.845;
    case 29: return 63.546;
    case 74: return 183.84;
    case 82: return 207.217;
  }
  return 2.0 * Z;
}

const MatDef* FindMaterial(const char* name) {
  for (const MatDef& m : kMaterials)
    if (std::strcmp(m.name, name) == 0) return &m;
  return nullptr;
}

// ---- numerics -----------------------------------------------------------------
// Gauss-Legendre nodes/weights on [0,1] (Newton iteration on P_n).
struct GL {
  std::vector<double> x, w;
  explicit GL(int n) : x(n), w(n) {
    for (int i = 0; i < (n + 1) / 2; ++i) {
      double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
      double pp = 0.0;
      for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = 0.0;
        for (int j = 1; j <= n; ++j) {
          const double p2 = p1;
          p1 = p0;
          p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
        }
        pp = n * (z * p0 - p1) / (z * z - 1.0);
        const double dz = p0 / pp;
        z -= dz;
        if (std::fabs(dz) < 1e-15) break;
      }
      x[i] = 0.5 * (1.0 - z);
      x[n - 1 - i] = 0.5 * (1.0 + z);
      w[i] = w[n - 1 - i] = 1.0 / ((1.0 - z * z) * pp * pp);
    }
  }
};
const GL& GL16() { static GL g(16); return g; }

// integral of f over [a,b] with `panels` 16-point Gauss-Legendre panels
template <class F>
double Integrate(F f, double a, double b, int panels) {
  const GL& g = GL16();
  const double h = (b - a) / panels;
  double s = 0.0;
  for (int p = 0; p < panels; ++p) {
    const double lo = a + p * h;
    double ps = 0.0;
    for (size_t j = 0; j < g.x.size(); ++j) ps += g.w[j] * f(lo + h * g.x[j]);
    s += ps * h;
  }
  return s;
}

// log-spaced grid with exact end points
void LogGrid(double emin, double emax, int n, double& logMin, double& invLDelta, double* grid) {
  const double delta = std::log(emax / emin) / (n - 1);
  logMin = std::log(emin);
  invLDelta = 1.0 / delta;
  grid[0] = emin;
  grid[n - 1] = emax;
  for (int i = 1; i < n - 1; ++i) grid[i] = std::exp(logMin + i * delta);
}

// natural cubic spline second derivatives d2y/dx2 (x strictly increasing)
void SplineSD(int n, const double* x, const double* y, double* sd, int ystride = 1) {
  std::vector<double> c(n, 0.0), d(n, 0.0);
  sd[0] = 0.0;
  sd[(n - 1) * ystride] = 0.0;
  if (n < 3) { for (int i = 0; i < n; ++i) sd[i * ystride] = 0.0; return; }
  // tridiagonal system for interior points (Thomas algorithm)
  for (int i = 1; i < n - 1; ++i) {
    const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
    const double a = h0 / 6.0, b = (h0 + h1) / 3.0, cc = h1 / 6.0;
    const double r = (y[(i + 1) * ystride] - y[i * ystride]) / h1 - (y[i * ystride] - y[(i - 1) * ystride]) / h0;
    const double den = b - a * c[i - 1];
    c[i] = cc / den;
    d[i] = (r - a * d[i - 1]) / den;
  }
  std::vector<double> m(n, 0.0);
  for (int i = n - 2; i >= 1; --i) m[i] = d[i] - c[i] * m[i + 1];
  for (int i = 0; i < n; ++i) sd[i * ystride] = m[i];
}

// ---- element / material models --------------------------------------------
double CoulombCorrection(int Z) {  // Davies-Bethe-Maximon f_c(Z)
  const double a2 = (kAlpha * Z) * (kAlpha * Z);
  return a2 * (1.0 / (1.0 + a2) + 0.20206 - 0.0369 * a2 + 0.0083 * a2 * a2 - 0.002 * a2 * a2 * a2);
}
// Tsai radiation logarithms (Tsai, Rev. Mod. Phys. 46 (1974) 815, Table B.2 for Z<5)
double Lrad(int Z)  { const double t[] = {0, 5.31, 4.79, 4.74, 4.71}; return Z < 5 ? t[Z] : std::log(184.15 * std::pow(Z, -1.0 / 3.0)); }
double Lprad(int Z) { const double t[] = {0, 6.144, 5.621, 5.805, 5.924}; return Z < 5 ? t[Z] : std::log(1194.0 * std::pow(Z, -2.0 / 3.0)); }
double KShellBinding(int Z) { return Z > 1 ? 13.6 * kEV * (Z - 1.0) * (Z - 1.0) : 13.6 * kEV; }
double LShellBinding(int Z) { return Z >= 10 ? 13.6 * kEV * (Z - 7.4) * (Z - 7.4) / 4.0 : 0.0; }

struct Mat {
  const MatDef* def;
  int nel;
  std::vector<int> Z;
  std::vector<double> nat;   // atoms / mm^3
  double nel_dens;           // electrons / mm^3
  double I;                  // MeV
  double X0;                 // mm
  double zeff;
  // Sternheimer-Peierls density-effect parameters
  double cbar, x0, x1, aSt;
};

Mat MakeMat(const MatDef* d) {
  Mat m;
  m.def = d;
  m.nel = d->nel;
  double molMass = 0.0;
  for (int i = 0; i < d->nel; ++i) molMass += d->count[i] * AtomicMass(d->Z[i]);
  m.nel_dens = 0.0;
  double invX0 = 0.0, sz2 = 0.0, sz = 0.0;
  for (int i = 0; i < d->nel; ++i) {
    const int Z = d->Z[i];
    const double n = d->rho * kNA * d->count[i] / molMass * 1.0e-3;  // 1/cm3 -> 1/mm3
    m.Z.push_back(Z);
    m.nat.push_back(n);
    m.nel_dens += n * Z;
    invX0 += n * 4.0 * kAlpha * kRe * kRe * (Z * Z * (Lrad(Z) - CoulombCorrection(Z)) + Z * Lprad(Z));
    sz2 += n * Z * Z;
    sz += n * Z;
  }
  m.X0 = 1.0 / invX0;
  m.zeff = sz2 / sz;  // electron-weighted mean Z
  m.I = d->meanExc * kEV;
  const double plasmaE = kHbarc * std::sqrt(4.0 * kPi * m.nel_dens * kRe);
  m.cbar = 1.0 + 2.0 * std::log(m.I / plasmaE);
  if (d->meanExc < 100.0) { m.x1 = 2.0; m.x0 = m.cbar < 3.681 ? 0.2 : 0.326 * m.cbar - 1.0; }
  else                    { m.x1 = 3.0; m.x0 = m.cbar < 5.215 ? 0.2 : 0.326 * m.cbar - 1.5; }
  m.aSt = (m.cbar - 4.606 * m.x0) / std::pow(m.x1 - m.x0, 3.0);
  return m;
}

double DensityEffect(const Mat& m, double bg2) {
  const double X = std::log(bg2) / (2.0 * std::log(10.0));
  if (X < m.x0) return 0.0;
  if (X < m.x1) return 4.606 * X - m.cbar + m.aSt * std::pow(m.x1 - X, 3.0);
  return 4.606 * X - m.cbar;
}

// Restricted collision stopping power (Berger-Seltzer, ICRU 37) [MeV/mm]
double DEDXIoniRaw(const Mat& m, double T, double cut, bool isEl) {
  const double tau = T / kMe, gam = tau + 1.0, gam2 = gam * gam;
  const double bg2 = tau * (tau + 2.0), beta2 = bg2 / gam2;
  const double eexc = m.I / kMe, eexc2 = eexc * eexc;
  const double tmax = isEl ? 0.5 * T : T;
  const double d = std::min(cut, tmax) / kMe;
  double f;
  if (isEl) {
    f = std::log(2.0 * (tau + 2.0) / eexc2) - 1.0 - beta2 + std::log((tau - d) * d) + tau / (tau - d)
        + (0.5 * d * d + (2.0 * tau + 1.0) * std::log(1.0 - d / tau)) / gam2;
  } else {
    const double d2 = d * d * 0.5, d3 = d2 * d / 1.5, d4 = d3 * d * 0.75;
    const double y = 1.0 / (1.0 + gam);
    f = std::log(2.0 * (tau + 2.0) / eexc2) + std::log(tau * d)
        - beta2 * (tau + 2.0 * d - y * (3.0 * d2 + y * (d - d3 + y * (d2 - tau * d3 + d4)))) / tau;
  }
  f -= DensityEffect(m, bg2);
  return std::max(0.0, 2.0 * kPi * kMe * kRe * kRe * m.nel_dens * f / beta2);
}
double DEDXIoni(const Mat& m, double T, double cut, bool isEl) {
  const double th = 0.25 * std::sqrt(m.zeff) * kKeV;
  if (T >= th) return DEDXIoniRaw(m, T, cut, isEl);
  const double x = T / th, d0 = DEDXIoniRaw(m, th, cut, isEl);
  return x > 0.25 ? d0 / std::sqrt(x) : d0 * 1.4 * std::sqrt(x) / (0.1 + x);
}

// Moller / Bhabha cross section per electron above `cut` [mm^2]
double XSIoniPerElectron(double T, double cut, bool isEl) {
  const double tmax = isEl ? 0.5 * T : T;
  if (cut >= tmax) return 0.0;
  const double xmin = cut / T, xmax = tmax / T;
  const double tau = T / kMe, gam = tau + 1.0, gam2 = gam * gam;
  const double beta2 = tau * (tau + 2.0) / gam2;
  double c;
  if (isEl) {
    const double gg = (2.0 * gam - 1.0) / gam2;
    c = ((xmax - xmin) * (1.0 - gg + 1.0 / (xmin * xmax) + 1.0 / ((1.0 - xmin) * (1.0 - xmax)))
         - gg * std::log(xmax * (1.0 - xmin) / (xmin * (1.0 - xmax)))) / beta2;
  } else {
    const double y = 1.0 / (1.0 + gam), y2 = y * y, y12 = 1.0 - 2.0 * y;
    const double b1 = 2.0 - y2, b2 = y12 * (3.0 + y2), y122 = y12 * y12, b4 = y122 * y12, b3 = b4 + y122;
    c = (xmax - xmin) * (1.0 / (beta2 * xmin * xmax) + b2 - 0.5 * b3 * (xmin + xmax)
                         + b4 * (xmin * xmin + xmin * xmax + xmax * xmax) / 3.0)
        - b1 * std::log(xmax / xmin);
  }
  return std::max(0.0, c * 2.0 * kPi * kMe * kRe * kRe / T);
}

// Bremsstrahlung: k dsigma/dk per atom [mm^2], Tsai complete screening
double KDCSBrem(int Z, double T, double k) {
  const double E = T + kMe, y = k / E;
  if (k >= T) return 0.0;
  const double fz = Z * Z * (Lrad(Z) - CoulombCorrection(Z)) + Z * Lprad(Z);
  return 4.0 * kAlpha * kRe * kRe * ((4.0 / 3.0 - 4.0 / 3.0 * y + y * y) * fz + (1.0 - y) * (Z * Z + Z) / 9.0);
}
// dielectric suppression factor k^2/(k^2+k_p^2) with k_p^2 = Migdal*n_e*E_tot^2
double DielSup(const Mat& m, double T, double k) {
  const double kp2 = kMigdal * m.nel_dens * (T + kMe) * (T + kMe);
  return k * k / (k * k + kp2);
}
// restricted brem cross section per atom for k in [cut, T] [mm^2]
double XSBremPerAtom(const Mat& m, int Z, double T, double cut) {
  if (cut >= T) return 0.0;
  return Integrate([&](double u) { const double k = std::exp(u); return KDCSBrem(Z, T, k) * DielSup(m, T, k); },
                   std::log(cut), std::log(T), 8);
}
double XSBrem(const Mat& m, double T, double cut) {
  double s = 0.0;
  for (int i = 0; i < m.nel; ++i) s += m.nat[i] * XSBremPerAtom(m, m.Z[i], T, cut);
  return s;
}
// restricted radiative stopping power: photons below cut [MeV/mm]
double DEDXBrem(const Mat& m, double T, double cut) {
  const double kup = std::min(cut, T);
  const double klo = kup * 1.0e-9;
  double s = 0.0;
  for (int i = 0; i < m.nel; ++i) {
    const int Z = m.Z[i];
    s += m.nat[i] * Integrate([&](double u) { const double k = std::exp(u); return k * KDCSBrem(Z, T, k) * DielSup(m, T, k); },
                              std::log(klo), std::log(kup), 8);
  }
  return s;
}

// screened-Rutherford first transport cross section per atom [mm^2]
double XSTr1PerAtom(int Z, double T) {
  const double pc2 = T * (T + 2.0 * kMe), E = T + kMe, beta2 = pc2 / (E * E);
  const double aTF = 0.88534 * kBohr * std::pow(Z, -1.0 / 3.0);
  const double A = (kHbarc * kHbarc / (4.0 * pc2 * aTF * aTF)) * (1.13 + 3.76 * kAlpha * kAlpha * Z * Z / beta2);
  const double K = Z * (Z + 1.0) * kRe * kRe * kMe * kMe * E * E / (pc2 * pc2);
  return 2.0 * kPi * K * (std::log1p(1.0 / A) - 1.0 / (1.0 + A));
}

// Klein-Nishina cross section per electron [mm^2]
double XSKleinNishina(double k) {
  const double x = k / kMe;
  const double sT = 8.0 * kPi * kRe * kRe / 3.0;
  if (x < 1.0e-3) return sT * (1.0 - 2.0 * x + 5.2 * x * x);
  const double l = std::log(1.0 + 2.0 * x);
  return 2.0 * kPi * kRe * kRe * ((1.0 + x) / (x * x) * (2.0 * (1.0 + x) / (1.0 + 2.0 * x) - l / x)
                                 + l / (2.0 * x) - (1.0 + 3.0 * x) / ((1.0 + 2.0 * x) * (1.0 + 2.0 * x)));
}
// Bethe-Heitler pair production per atom [mm^2] (complete screening, Coulomb
// corrected, smooth threshold factor)
double XSPairPerAtom(int Z, double k) {
  if (k <= 2.0 * kMe) return 0.0;
  const double sInf = 4.0 * kAlpha * kRe * kRe * Z * (Z + 1.0)
                      * (7.0 / 9.0 * (std::log(183.0 * std::pow(Z, -1.0 / 3.0)) - CoulombCorrection(Z)) - 1.0 / 54.0);
  const double t = 1.0 - 2.0 * kMe / k;
  return sInf * t * t * t;
}

// Photoabsorption per atom in the 4-coefficient Sandia form
// sigma(E) = a1/E + a2/E^2 + a3/E^3 + a4/E^4 on intervals starting at
// 10 eV, the L edge (Z>=10) and the K edge. Coefficients [mm^2 MeV^k].
struct Sandia {
  std::vector<double> e;     // interval lower edges
  std::vector<double> c;     // 4 per interval
};
Sandia ElementSandia(int Z) {
  Sandia s;
  const double norm = 1800.0 * kBarn * std::pow(Z / 82.0, 4.5);
  const double a1 = norm * 0.003 * 0.1, a3 = norm * 0.1 * 0.1 * 0.1;
  const double eK = KShellBinding(Z), eL = LShellBinding(Z);
  s.e.push_back(10.0 * kEV);
  if (eL > 10.0 * kEV) {
    s.c.insert(s.c.end(), {a1 / 20.0, 0.0, a3 / 20.0, 0.0});
    s.e.push_back(eL);
  }
  s.c.insert(s.c.end(), {a1 / 5.0, 0.0, a3 / 5.0, 0.0});
  s.e.push_back(eK);
  s.c.insert(s.c.end(), {a1, 0.0, a3, 0.0});
  return s;
}
double SandiaEval(const Sandia& s, double E) {
  int iv = 0;
  for (int i = (int)s.e.size() - 1; i >= 0; --i)
    if (E >= s.e[i]) { iv = i; break; }
  const double* c = &s.c[4 * iv];
  const double inv = 1.0 / E;
  return inv * (c[0] + inv * (c[1] + inv * (c[2] + inv * c[3])));
}
Sandia MaterialSandia(const Mat& m) {
  std::vector<double> edges;
  for (int i = 0; i < m.nel; ++i) {
    Sandia es = ElementSandia(m.Z[i]);
    edges.insert(edges.end(), es.e.begin(), es.e.end());
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  Sandia s;
  s.e = edges;
  s.c.assign(4 * edges.size(), 0.0);
  for (int i = 0; i < m.nel; ++i) {
    Sandia es = ElementSandia(m.Z[i]);
    for (size_t iv = 0; iv < edges.size(); ++iv) {
      int j = 0;
      for (int jj = (int)es.e.size() - 1; jj >= 0; --jj)
        if (edges[iv] >= es.e[jj]) { j = jj; break; }
      for (int c = 0; c < 4; ++c) s.c[4 * iv + c] += m.nat[i] * es.c[4 * j + c];
    }
  }
  return s;
}

// ---- production-cut conversion (range cut -> energy) ------------------------
double RangeCutToEnergyLepton(const Mat& m, double cutMM, bool isEl) {
  const int n = 400;
  const double e0 = 1.0 * kKeV, e1 = 10.0e3;
  std::vector<double> E(n), R(n);
  double lmin, ild;
  LogGrid(e0, e1, n, lmin, ild, E.data());
  auto S = [&](double T) { return DEDXIoni(m, T, T, isEl) + DEDXBrem(m, T, T); };
  R[0] = 2.0 * E[0] / S(E[0]);
  for (int i = 1; i < n; ++i)
    R[i] = R[i - 1] + Integrate([&](double u) { const double T = std::exp(u); return T / S(T); },
                                std::log(E[i - 1]), std::log(E[i]), 1);
  if (cutMM <= R[0]) return std::max(990.0 * kEV, E[0] * (cutMM / R[0]) * (cutMM / R[0]));
  for (int i = 1; i < n; ++i)
    if (R[i] >= cutMM) {
      const double f = (cutMM - R[i - 1]) / (R[i] - R[i - 1]);
      return std::max(990.0 * kEV, std::exp(std::log(E[i - 1]) + f * (std::log(E[i]) - std::log(E[i - 1]))));
    }
  return e1;
}
double RangeCutToEnergyGamma(const Mat& m, double cutMM) {
  const Sandia sm = MaterialSandia(m);
  auto sigma = [&](double k) {
    double s = SandiaEval(sm, k) + m.nel_dens * XSKleinNishina(k);
    for (int i = 0; i < m.nel; ++i) s += m.nat[i] * XSPairPerAtom(m.Z[i], k);
    return s;
  };
  const double target = 5.0 / cutMM;
  const double lo = std::log(990.0 * kEV), hi = std::log(10.0e3);
  const int n = 2000;
  double prev = lo;
  if (sigma(std::exp(lo)) < target) return 990.0 * kEV;
  for (int i = 1; i <= n; ++i) {
    const double u = lo + (hi - lo) * i / n;
    if (sigma(std::exp(u)) < target) {
      double a = prev, b = u;
      for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (a + b);
        if (sigma(std::exp(mid)) < target) b = mid; else a = mid;
      }
      return std::exp(0.5 * (a + b));
    }
    prev = u;
  }
  return 10.0e3;
}

// ---- allocation helpers ------------------------------------------------------
template <class T> T* NewArr(size_t n) { return n ? static_cast<T*>(std::calloc(n, sizeof(T))) : nullptr; }
template <class T> T* CopyArr(const std::vector<T>& v) {
  T* p = NewArr<T>(v.size());
  if (p) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

// ---- builders ------------------------------------------------------------------
struct Couple { double eCut, pCut, gCut; int imat; };

void BuildElectronData(G4HepEmElectronData* ed, const std::vector<Mat>& mats, const std::vector<Couple>& cps,
                       const G4HepEmParameters* pars, bool isEl) {
  const int nmc = (int)cps.size(), nmat = (int)mats.size();
  ed->fNumMatCuts = nmc;
  ed->fNumMaterials = nmat;
  // -- energy loss
  const int N = pars->fNumLossTableBins + 1;
  ed->fELossEnergyGridSize = N;
  ed->fELossEnergyGrid = NewArr<double>(N);
  LogGrid(pars->fMinLossTableEnergy, pars->fMaxLossTableEnergy, N, ed->fELossLogMinEkin, ed->fELossEILDelta, ed->fELossEnergyGrid);
  const double* eg = ed->fELossEnergyGrid;
  ed->fELossData = NewArr<double>(5 * N * nmc);
  std::vector<double> dedx(N), rng(N), sdD(N), sdR(N), sdI(N);
  for (int imc = 0; imc < nmc; ++imc) {
    const Mat& m = mats[cps[imc].imat];
    const double eCut = cps[imc].eCut, gCut = cps[imc].gCut;
    auto S = [&](double T) { return DEDXIoni(m, T, eCut, isEl) + DEDXBrem(m, T, gCut); };
    for (int i = 0; i < N; ++i) dedx[i] = S(eg[i]);
    // range: R(E_0) = 2 E_0 / S(E_0) (S ~ sqrt(E) below E_0), then integrate 1/S
    rng[0] = 2.0 * eg[0] / dedx[0];
    for (int i = 0; i < N - 1; ++i)
      rng[i + 1] = rng[i] + Integrate([&](double u) { const double T = std::exp(u); return T / S(T); },
                                      std::log(eg[i]), std::log(eg[i + 1]), 1);
    SplineSD(N, eg, dedx.data(), sdD.data());
    SplineSD(N, eg, rng.data(), sdR.data());
    SplineSD(N, rng.data(), eg, sdI.data());
    double* d = &ed->fELossData[5 * N * imc];
    for (int i = 0; i < N; ++i) {
      d[2 * i] = rng[i];
      d[2 * i + 1] = sdR[i];
      d[2 * (N + i)] = dedx[i];
      d[2 * (N + i) + 1] = sdD[i];
      d[4 * N + i] = sdI[i];
    }
  }
  // -- restricted macroscopic cross sections (ioni then brem per couple)
  std::vector<double> xs;
  ed->fResMacXSecStartIndexPerMatCut = NewArr<int>(nmc);
  const double emax = pars->fMaxLossTableEnergy;
  const double scale = std::log(pars->fMaxLossTableEnergy / pars->fMinLossTableEnergy);
  for (int imc = 0; imc < nmc; ++imc) {
    const Mat& m = mats[cps[imc].imat];
    ed->fResMacXSecStartIndexPerMatCut[imc] = (int)xs.size();
    for (int ip = 0; ip < 2; ++ip) {
      const double emin = ip == 0 ? (isEl ? 2.0 * cps[imc].eCut : cps[imc].eCut) : cps[imc].gCut;
      const int M = std::max(4, (int)std::lrint(N * std::log(emax / emin) / scale) + 1);
      std::vector<double> e(M), s(M), sd(M);
      double lmin, ild;
      LogGrid(emin, emax, M, lmin, ild, e.data());
      double smax = -1.0, emaxv = -1.0;
      for (int i = 0; i < M; ++i) {
        s[i] = ip == 0 ? m.nel_dens * XSIoniPerElectron(e[i], cps[imc].eCut, isEl) : XSBrem(m, e[i], cps[imc].gCut);
        if (s[i] > smax) { smax = s[i]; emaxv = e[i]; }
      }
      SplineSD(M, e.data(), s.data(), sd.data());
      xs.insert(xs.end(), {(double)M, emaxv, smax, lmin, ild});
      for (int i = 0; i < M; ++i) xs.insert(xs.end(), {e[i], s[i], sd[i]});
    }
  }
  ed->fResMacXSecNumData = (int)xs.size();
  ed->fResMacXSecData = CopyArr(xs);
  // -- e-/e+ nuclear: 128-point grid 100 MeV - 100 TeV, no hadronic data (zero)
  ed->fENucEnergyGridSize = 128;
  ed->fENucEnergyGrid = NewArr<double>(128);
  LogGrid(100.0, 1.0e8, 128, ed->fENucLogMinEkin, ed->fENucEILDelta, ed->fENucEnergyGrid);
  ed->fENucMacXsecData = NewArr<double>(2 * 128 * nmat);
  // -- first transport cross section per material on the loss grid
  ed->fTr1MacXSecData = NewArr<double>(2 * N * nmat);
  std::vector<double> tr(N), sdT(N);
  for (int im = 0; im < nmat; ++im) {
    const Mat& m = mats[im];
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
      for (int k = 0; k < m.nel; ++k) s += m.nat[k] * XSTr1PerAtom(m.Z[k], eg[i]);
      tr[i] = s;
    }
    SplineSD(N, eg, tr.data(), sdT.data());
    for (int i = 0; i < N; ++i) {
      ed->fTr1MacXSecData[2 * N * im + 2 * i] = tr[i];
      ed->fTr1MacXSecData[2 * N * im + 2 * i + 1] = sdT[i];
    }
  }
  // -- element selectors (ioni, SB brem, RB brem)
  const int binsPerDecade = 7;
  auto buildSel = [&](int which, int* startIdx, int& numData, double*& data) {
    std::vector<double> v;
    for (int imc = 0; imc < nmc; ++imc) {
      const Mat& m = mats[cps[imc].imat];
      startIdx[imc] = -1;
      if (m.nel < 2) continue;
      double emin, emx;
      if (which == 0)      { emin = isEl ? 2.0 * cps[imc].eCut : cps[imc].eCut; emx = pars->fMaxLossTableEnergy; }
      else if (which == 1) { emin = cps[imc].gCut; emx = pars->fElectronBremModelLim; }
      else                 { emin = std::max(cps[imc].gCut, pars->fElectronBremModelLim); emx = pars->fMaxLossTableEnergy; }
      if (emin >= emx) continue;
      int K = (int)(binsPerDecade * std::log(emx / emin) / (6.0 * std::log(10.0)));
      K = std::max(K, 3) + 1;
      std::vector<double> e(K);
      double lmin, ild;
      LogGrid(emin, emx, K, lmin, ild, e.data());
      startIdx[imc] = (int)v.size();
      v.insert(v.end(), {(double)K, (double)m.nel, lmin, ild});
      for (int i = 0; i < K; ++i) {
        v.push_back(e[i]);
        const size_t ist = v.size();
        double sum = 0.0;
        for (int iz = 0; iz < m.nel; ++iz) {
          const double x = which == 0 ? m.Z[iz] * XSIoniPerElectron(e[i], cps[imc].eCut, isEl)
                                      : XSBremPerAtom(m, m.Z[iz], e[i], cps[imc].gCut);
          sum += m.nat[iz] * x;
          if (iz < m.nel - 1) v.push_back(sum);
        }
        for (int iz = 0; iz < m.nel - 1; ++iz)
          if (sum > 0.0) v[ist + iz] /= sum;
      }
    }
    numData = (int)v.size();
    data = CopyArr(v);
  };
  ed->fElemSelectorIoniStartIndexPerMatCut = NewArr<int>(nmc);
  ed->fElemSelectorBremSBStartIndexPerMatCut = NewArr<int>(nmc);
  ed->fElemSelectorBremRBStartIndexPerMatCut = NewArr<int>(nmc);
  buildSel(0, ed->fElemSelectorIoniStartIndexPerMatCut, ed->fElemSelectorIoniNumData, ed->fElemSelectorIoniData);
  buildSel(1, ed->fElemSelectorBremSBStartIndexPerMatCut, ed->fElemSelectorBremSBNumData, ed->fElemSelectorBremSBData);
  buildSel(2, ed->fElemSelectorBremRBStartIndexPerMatCut, ed->fElemSelectorBremRBNumData, ed->fElemSelectorBremRBData);
}

// RITA (rational inverse transform with aliasing) parameters of an interval
// (Salvat et al., PENELOPE manual, sect. 1.2.2): for a pdf p(x) with cumulative
// xi(x) on [x_j, x_{j+1}]: b = 1 - (dxi/dx)^2/(p_j p_{j+1}), a = dxi/(dx p_j) - b - 1.
void BuildSBTables(G4HepEmSBTableData* sb, const G4HepEmData* hd, const std::vector<Mat>& mats,
                   const std::vector<Couple>& cps, const G4HepEmParameters* pars) {
  sb->fMaxZet = 99;
  sb->fNumElEnergy = 65;
  sb->fNumKappa = 54;
  const double elEmin = 100.0 * kEV, elEmax = 10.0e3;
  sb->fLogMinElEnergy = std::log(elEmin);
  sb->fILDeltaElEnergy = 1.0 / (std::log(elEmax / elEmin) / (65 - 1.0));
  for (int i = 0; i < 65; ++i) {
    sb->fElEnergyVect[i] = i == 0 ? elEmin : (i == 64 ? elEmax : std::exp(sb->fLogMinElEnergy + i / sb->fILDeltaElEnergy));
    sb->fLElEnergyVect[i] = std::log(sb->fElEnergyVect[i]);
  }
  // kappa grid: 30 points log-spaced on [1e-12, 1e-2), 24 on [1e-2, 1]
  for (int i = 0; i < 54; ++i) {
    double lk;
    if (i < 30) lk = std::log(1.0e-12) + i * (std::log(1.0e-2) - std::log(1.0e-12)) / 30.0;
    else        lk = std::log(1.0e-2) + (i - 30) * (0.0 - std::log(1.0e-2)) / 23.0;
    sb->fKappaVect[i] = i == 53 ? 1.0 : std::exp(lk);
    sb->fLKappaVect[i] = i == 53 ? 0.0 : lk;
  }
  const int nmc = (int)cps.size();
  sb->fNumHepEmMatCuts = nmc;
  int numElemsInMC = 0;
  for (int imc = 0; imc < nmc; ++imc) numElemsInMC += mats[cps[imc].imat].nel;
  sb->fNumElemsInMatCuts = numElemsInMC;
  sb->fGammaCutIndxStartIndexPerMC = NewArr<int>(nmc);
  sb->fGammaCutIndices = NewArr<int>(numElemsInMC);
  for (int i = 0; i < 121; ++i) sb->fSBTablesStartPerZ[i] = 0;

  const double usedLow = std::max(pars->fElectronTrackingCut, 1.0 * kKeV);
  const double usedHigh = std::min(pars->fElectronBremModelLim, elEmax);
  std::vector<double> data;
  // per Z: sorted distinct gamma cuts (below the SB upper limit) and couple -> cut index
  std::vector<std::vector<int>> cutIndexOfCouple(121, std::vector<int>(nmc, -1));
  const G4HepEmElementData* elD = hd->fTheElementData;
  for (int iz = 1; iz < elD->fMaxZet; ++iz) {
    if (elD->fElementData[iz].fZet < 0.0) continue;
    std::vector<double> cuts;
    for (int imc = 0; imc < nmc; ++imc) {
      const Mat& m = mats[cps[imc].imat];
      if (std::find(m.Z.begin(), m.Z.end(), iz) == m.Z.end()) continue;
      if (cps[imc].gCut >= usedHigh) continue;
      if (std::find(cuts.begin(), cuts.end(), cps[imc].gCut) == cuts.end()) cuts.push_back(cps[imc].gCut);
    }
    if (cuts.empty()) continue;
    std::sort(cuts.begin(), cuts.end());
    for (int imc = 0; imc < nmc; ++imc) {
      const Mat& m = mats[cps[imc].imat];
      if (std::find(m.Z.begin(), m.Z.end(), iz) == m.Z.end()) continue;
      const auto it = std::find(cuts.begin(), cuts.end(), cps[imc].gCut);
      if (it != cuts.end()) cutIndexOfCouple[iz][imc] = (int)(it - cuts.begin());
    }
    const double lowE = std::max(usedLow, cuts[0]);
    const double* ev = sb->fElEnergyVect;
    const int minIdx = (lowE >= ev[64]) ? 64 : (int)(std::lower_bound(ev, ev + 65, lowE) - ev) - 1;
    const int maxIdx = (usedHigh >= ev[64]) ? 64 : (int)(std::lower_bound(ev, ev + 65, usedHigh) - ev);
    const int ncut = (int)cuts.size();
    sb->fSBTablesStartPerZ[iz] = (int)data.size();
    data.insert(data.end(), {(double)((maxIdx - minIdx + 1) * (ncut + 3 * 54) + 4), (double)minIdx, (double)maxIdx, (double)ncut});
    for (int ie = minIdx; ie <= maxIdx; ++ie) {
      const double T = ev[ie];
      // pdf in x = ln(kappa): k dsigma/dk at k = kappa T (no dielectric suppression)
      auto pdf = [&](double x) { return KDCSBrem(iz, T, std::exp(x) * T * (1.0 - 1.0e-12)); };
      double cum[54], p[54], A[54], B[54];
      cum[0] = 0.0;
      for (int j = 0; j < 53; ++j)
        cum[j + 1] = cum[j] + Integrate(pdf, sb->fLKappaVect[j], sb->fLKappaVect[j + 1], 2);
      const double norm = cum[53];
      for (int j = 0; j < 54; ++j) {
        cum[j] /= norm;
        p[j] = pdf(sb->fLKappaVect[j]) / norm;
      }
      cum[53] = 1.0;
      for (int j = 0; j < 53; ++j) {
        const double dx = sb->fLKappaVect[j + 1] - sb->fLKappaVect[j];
        const double dxi = cum[j + 1] - cum[j];
        B[j] = 1.0 - (dxi / dx) * (dxi / dx) / (p[j] * p[j + 1]);
        A[j] = dxi / (dx * p[j]) - B[j] - 1.0;
      }
      A[53] = B[53] = 0.0;
      // cumulative value at each gamma cut (1 when no emission above the cut is possible)
      for (int ic = 0; ic < ncut; ++ic) {
        double val = 1.0;
        if (T > cuts[ic]) {
          const double ck = std::max(1.0e-12, cuts[ic] / T);
          const int jl = ck > 1.0e-12 ? std::max(0, (int)(std::lower_bound(sb->fKappaVect, sb->fKappaVect + 54, ck) - sb->fKappaVect) - 1) : 0;
          const double alph = std::log(ck / sb->fKappaVect[jl]) / (sb->fLKappaVect[jl + 1] - sb->fLKappaVect[jl]);
          // invert alph = (1+A+B) eta / (1 + A eta + B eta^2) for eta in [0,1]
          const double a = A[jl], b = B[jl];
          double eta;
          const double qa = b * alph, qb = a * alph - 1.0 - a - b, qc = alph;
          if (std::fabs(qa) < 1.0e-14) eta = -qc / qb;
          else eta = (-qb - std::sqrt(std::max(0.0, qb * qb - 4.0 * qa * qc))) / (2.0 * qa);
          eta = std::min(1.0, std::max(0.0, eta));
          val = cum[jl] + eta * (cum[jl + 1] - cum[jl]);
        }
        data.push_back(val);
      }
      for (int j = 0; j < 54; ++j) data.insert(data.end(), {cum[j], A[j], B[j]});
    }
  }
  sb->fNumSBTableData = (int)data.size();
  sb->fSBTableData = CopyArr(data);
  int ind = 0;
  for (int imc = 0; imc < nmc; ++imc) {
    const Mat& m = mats[cps[imc].imat];
    bool first = true;
    sb->fGammaCutIndxStartIndexPerMC[imc] = 0;
    for (int ie = 0; ie < m.nel; ++ie) {
      const int z = std::min(m.Z[ie], 120);
      const int ci = cutIndexOfCouple[z][imc];
      if (ci < 0) continue;
      if (first) { sb->fGammaCutIndxStartIndexPerMC[imc] = ind; first = false; }
      sb->fGammaCutIndices[ind++] = ci;
    }
  }
}

void BuildGammaData(G4HepEmGammaData* gd, const std::vector<Mat>& mats) {
  const int nmat = (int)mats.size();
  gd->fNumMaterials = nmat;
  gd->fEGridSize0 = 32;
  gd->fEGridSize1 = 32;
  gd->fEGridSize2 = 256;
  const int n0 = 32, n1 = 32, n2 = 256;
  std::vector<double> g0(n0), g1(n1), g2(n2);
  gd->fEMin0 = 100.0 * kEV;
  gd->fEMax0 = 150.0 * kKeV;
  LogGrid(gd->fEMin0, gd->fEMax0, n0, gd->fLogEMin0, gd->fEILDelta0, g0.data());
  gd->fEMax1 = 2.0 * kMe;
  LogGrid(gd->fEMax0, gd->fEMax1, n1, gd->fLogEMin1, gd->fEILDelta1, g1.data());
  gd->fEMax2 = 100.0e6;
  LogGrid(gd->fEMax1, gd->fEMax2, n2, gd->fLogEMin2, gd->fEILDelta2, g2.data());
  gd->fNumData0 = 2 * n0;
  gd->fNumData1 = 3 * n1;
  gd->fDataPerMat = 2 * n0 + 3 * n1 + 9 * n2;
  // 16 zero doubles of tail padding: the reference's gamma SampleInteraction
  // (G4HepEmGammaManager.icc:203-208) reads two values past a material's
  // block when it falls into the gamma-nuclear bin
  gd->fMacXsecData = NewArr<double>((size_t)nmat * gd->fDataPerMat + 16);
  for (int im = 0; im < nmat; ++im) {
    const Mat& m = mats[im];
    const Sandia sm = MaterialSandia(m);
    auto comp = [&](double k) { return m.nel_dens * XSKleinNishina(k); };
    auto pe = [&](double k) { return SandiaEval(sm, k); };
    auto conv = [&](double k) { double s = 0.0; for (int i = 0; i < m.nel; ++i) s += m.nat[i] * XSPairPerAtom(m.Z[i], k); return s; };
    double* d = &gd->fMacXsecData[(size_t)im * gd->fDataPerMat];
    for (int i = 0; i < n0; ++i) { d[2 * i] = g0[i]; d[2 * i + 1] = comp(g0[i]); }
    d += 2 * n0;
    for (int i = 0; i < n1; ++i) { d[3 * i] = g1[i]; d[3 * i + 1] = comp(g1[i]) + pe(g1[i]); d[3 * i + 2] = pe(g1[i]); }
    d += 3 * n1;
    std::vector<double> y[4], s[4];
    for (int q = 0; q < 4; ++q) { y[q].resize(n2); s[q].resize(n2); }
    for (int i = 0; i < n2; ++i) {
      y[1][i] = conv(g2[i]);
      y[2][i] = comp(g2[i]);
      y[3][i] = pe(g2[i]);
      y[0][i] = y[1][i] + y[2][i] + y[3][i];
    }
    for (int q = 0; q < 4; ++q) SplineSD(n2, g2.data(), y[q].data(), s[q].data());
    for (int i = 0; i < n2; ++i) {
      d[9 * i] = g2[i];
      for (int q = 0; q < 4; ++q) { d[9 * i + 1 + 2 * q] = y[q][i]; d[9 * i + 2 + 2 * q] = s[q][i]; }
    }
  }
  // conversion element selector
  const double emin = gd->fEMax1, emx = gd->fEMax2;
  const int nk = (int)(7 * std::log(emx / emin) / (6.0 * std::log(10.0)));
  gd->fElemSelectorConvEgridSize = nk;
  gd->fElemSelectorConvEgrid = NewArr<double>(nk);
  LogGrid(emin, emx, nk, gd->fElemSelectorConvLogMinEkin, gd->fElemSelectorConvEILDelta, gd->fElemSelectorConvEgrid);
  gd->fElemSelectorConvStartIndexPerMat = NewArr<int>(nmat);
  std::vector<double> v;
  for (int im = 0; im < nmat; ++im) {
    const Mat& m = mats[im];
    if (m.nel < 2) { gd->fElemSelectorConvStartIndexPerMat[im] = -1; continue; }
    gd->fElemSelectorConvStartIndexPerMat[im] = (int)v.size();
    v.push_back((double)m.nel);
    for (int ie = 0; ie < nk; ++ie) {
      const double k = gd->fElemSelectorConvEgrid[ie];
      const size_t ist = v.size();
      double sum = 0.0;
      for (int iz = 0; iz < m.nel; ++iz) {
        sum += m.nat[iz] * XSPairPerAtom(m.Z[iz], k);
        if (iz < m.nel - 1) v.push_back(sum);
      }
      if (sum > 0.0)
        for (int iz = 0; iz < m.nel - 1; ++iz) v[ist + iz] *= 1.0 / sum;
    }
  }
  gd->fElemSelectorConvNumData = (int)v.size();
  gd->fElemSelectorConvData = CopyArr(v);
}

// ---- serialisation -------------------------------------------------------------
struct Writer {
  uint8_t* buf;
  uint64_t pos = 0;
  template <class T> void put(const T& v) { if (buf) std::memcpy(buf + pos, &v, sizeof(T)); pos += sizeof(T); }
  template <class T> void arr(const T* p, int64_t n) {
    put<int64_t>(n);
    if (n > 0 && p) { if (buf) std::memcpy(buf + pos, p, n * sizeof(T)); pos += n * sizeof(T); }
  }
};

}  // namespace

extern "C" int hepem_tables_build(const char* const* names, int nmat, double cutMM,
                                  G4HepEmData** outData, G4HepEmParameters** outPars) {
  if (!names || nmat <= 0 || !outData || !outPars || !(cutMM > 0.0)) return HEPEM_ERR_INVALID;
  std::vector<Mat> mats;
  for (int i = 0; i < nmat; ++i) {
    if (!names[i]) return HEPEM_ERR_INVALID;
    const MatDef* d = FindMaterial(names[i]);
    if (!d) return HEPEM_ERR_INVALID;
    mats.push_back(MakeMat(d));
  }
  // parameters: G4EmParameters defaults, one region
  G4HepEmParameters* pars = NewArr<G4HepEmParameters>(1);
  pars->fElectronTrackingCut = 1.0 * kKeV;
  pars->fGammaTrackingCut = 0.0;
  pars->fMinLossTableEnergy = 100.0 * kEV;
  pars->fMaxLossTableEnergy = 100.0e6;
  pars->fNumLossTableBins = 84;
  pars->fElectronBremModelLim = 1.0e3;
  pars->fIsMSCPositronCor = true;
  pars->fIsMSCDisplacement = true;
  pars->fNumRegions = 1;
  pars->fParametersPerRegion = NewArr<G4HepEmRegionParmeters>(1);
  G4HepEmRegionParmeters& rp = pars->fParametersPerRegion[0];
  rp.fFinalRange = 1.0;
  rp.fDRoverRange = 0.2;
  rp.fLinELossLimit = 0.01;
  rp.fMSCRangeFactor = 0.04;
  rp.fMSCSafetyFactor = 0.6;
  rp.fIsMSCMinimalStepLimit = false;
  rp.fIsELossFluctuation = true;
  rp.fIsMultipleStepsInMSCTrans = true;
  rp.fIsApplyCuts = true;

  // couples: one per material, all with the same range cut
  std::vector<Couple> cps;
  for (int i = 0; i < nmat; ++i) {
    Couple c;
    c.imat = i;
    c.eCut = RangeCutToEnergyLepton(mats[i], cutMM, true);
    c.pCut = RangeCutToEnergyLepton(mats[i], cutMM, false);
    c.gCut = RangeCutToEnergyGamma(mats[i], cutMM);
    cps.push_back(c);
  }

  G4HepEmData* hd = NewArr<G4HepEmData>(1);
  // -- couples
  hd->fTheMatCutData = NewArr<G4HepEmMatCutData>(1);
  G4HepEmMatCutData* mcd = hd->fTheMatCutData;
  mcd->fNumG4MatCuts = nmat;
  mcd->fNumMatCutData = nmat;
  mcd->fG4MCIndexToHepEmMCIndex = NewArr<int>(nmat);
  mcd->fMatCutData = NewArr<G4HepEmMCCData>(nmat);
  for (int i = 0; i < nmat; ++i) {
    mcd->fG4MCIndexToHepEmMCIndex[i] = i;
    G4HepEmMCCData& c = mcd->fMatCutData[i];
    c.fSecElProdCutE = std::max(cps[i].eCut, pars->fElectronTrackingCut);
    c.fSecPosProdCutE = cps[i].pCut;
    c.fSecGamProdCutE = cps[i].gCut;
    c.fLogSecGamCutE = std::log(cps[i].gCut);
    c.fHepEmMatIndex = i;
    c.fG4MatCutIndex = i;
    c.fG4RegionIndex = 0;
    cps[i].eCut = c.fSecElProdCutE;
  }
  // -- elements (indexed by Z, 121 slots as the reference)
  hd->fTheElementData = NewArr<G4HepEmElementData>(1);
  G4HepEmElementData* eld = hd->fTheElementData;
  eld->fMaxZet = 120;
  eld->fElementData = NewArr<G4HepEmElemData>(121);
  for (int z = 0; z < 121; ++z) eld->fElementData[z].fZet = -1.0;
  // -- materials
  hd->fTheMaterialData = NewArr<G4HepEmMaterialData>(1);
  G4HepEmMaterialData* md = hd->fTheMaterialData;
  md->fNumG4Material = nmat;
  md->fNumMaterialData = nmat;
  md->fG4MatIndexToHepEmMatIndex = NewArr<int>(nmat);
  md->fMaterialData = NewArr<G4HepEmMatData>(nmat);
  for (int i = 0; i < nmat; ++i) {
    md->fG4MatIndexToHepEmMatIndex[i] = i;
    const Mat& m = mats[i];
    G4HepEmMatData& d = md->fMaterialData[i];
    d.fG4MatIndex = i;
    d.fNumOfElement = m.nel;
    d.fElementVect = CopyArr(m.Z);
    d.fNumOfAtomsPerVolumeVect = CopyArr(m.nat);
    d.fDensity = m.def->rho;
    d.fDensityCorFactor = kMigdal * m.nel_dens;
    d.fElectronDensity = m.nel_dens;
    d.fRadiationLength = m.X0;
    d.fMeanExEnergy = m.I;
    const Sandia s = MaterialSandia(m);
    d.fNumOfSandiaIntervals = (int)s.e.size();
    d.fSandiaEnergies = CopyArr(s.e);
    d.fSandiaCoefficients = CopyArr(s.c);
    // Urban MSC material parameters (Geant4 >= 10.7 parametrisation of the
    // Urban model; G4HepEmMaterialInit.cc:125-142 lists the same forms)
    const double zeff = m.zeff, z16 = std::pow(zeff, 1.0 / 6.0), z13 = z16 * z16;
    d.fZeff = zeff;
    d.fZeff23 = z13 * z13;
    d.fZeffSqrt = std::sqrt(zeff);
    d.fUMSCPar = 9.62800E-1 - 8.4848E-2 * d.fZeffSqrt + 4.3769E-3 * zeff;
    d.fUMSCStepMinPars[0] = 2.7725E+1 / (1.0 + 2.03E-1 * zeff);
    d.fUMSCStepMinPars[1] = 6.152 / (1.0 + 1.11E-1 * zeff);
    const double t0 = 9.90395E-1 + z16 * (-1.68386E-1 + z16 * 9.3286E-2);
    d.fUMSCThetaCoeff[0] = t0 * (1.0 - 8.7780E-2 / zeff);
    d.fUMSCThetaCoeff[1] = t0 * (4.0780E-2 + 1.7315E-4 * zeff);
    d.fUMSCTailCoeff[0] = 2.3785 - z13 * (4.1981E-1 - z13 * 6.3100E-2);
    d.fUMSCTailCoeff[1] = 4.7526E-1 + z13 * (1.7694 - z13 * 3.3885E-1);
    d.fUMSCTailCoeff[2] = 2.3683E-1 - z13 * (1.8111 - z13 * 3.2774E-1);
    d.fUMSCTailCoeff[3] = 1.7888E-2 + z13 * (1.9659E-2 - z13 * 2.6664E-3);
    for (int ie = 0; ie < m.nel; ++ie) {
      const int Z = std::min(m.Z[ie], 120);
      G4HepEmElemData& e = eld->fElementData[Z];
      if (e.fZet >= 0.0) continue;
      const double dZ = (double)Z;
      e.fZet = dZ;
      e.fZet13 = std::pow(dZ, 1.0 / 3.0);
      e.fZet23 = std::pow(dZ, 2.0 / 3.0);
      e.fCoulomb = CoulombCorrection(Z);
      e.fLogZ = std::log(dZ);
      // screening / LPM constants consumed by the RB-brem and conversion samplers
      e.fZFactor1 = (Lrad(Z) - e.fCoulomb) + Lprad(Z) / dZ;
      const double fzLow = 8.0 * e.fLogZ / 3.0, fzHigh = 8.0 * (e.fLogZ / 3.0 + e.fCoulomb);
      e.fDeltaMaxLow = std::exp((42.038 - fzLow) / 8.29) - 0.958;
      e.fDeltaMaxHigh = std::exp((42.038 - fzHigh) / 8.29) - 0.958;
      const double s1 = e.fZet23 / (184.15 * 184.15);
      e.fILVarS1Cond = 1.0 / std::log(std::sqrt(2.0) * s1);
      e.fILVarS1 = 1.0 / std::log(s1);
      const Sandia es = ElementSandia(Z);
      e.fNumOfSandiaIntervals = (int)es.e.size();
      e.fSandiaEnergies = CopyArr(es.e);
      e.fSandiaCoefficients = CopyArr(es.c);
      e.fKShellBindingEnergy = KShellBinding(Z);
    }
  }
  // -- e-/e+ tables
  hd->fTheElectronData = NewArr<G4HepEmElectronData>(1);
  BuildElectronData(hd->fTheElectronData, mats, cps, pars, true);
  hd->fThePositronData = NewArr<G4HepEmElectronData>(1);
  BuildElectronData(hd->fThePositronData, mats, cps, pars, false);
  // -- SB sampling tables and gamma data
  hd->fTheSBTableData = NewArr<G4HepEmSBTableData>(1);
  BuildSBTables(hd->fTheSBTableData, hd, mats, cps, pars);
  hd->fTheGammaData = NewArr<G4HepEmGammaData>(1);
  BuildGammaData(hd->fTheGammaData, mats);
  *outData = hd;
  *outPars = pars;
  return HEPEM_OK;
}

// Lepto- and photo-nuclear macroscopic cross sections in units of the
// inverse radiation length of the material, for tables that exercise the
// nuclear branches. e-/e+: slowly rising above the 100 MeV grid start.
// gamma: a giant-dipole bump at 13 MeV (threshold 5 MeV) over a plateau.
namespace {
double ENucShape(double E) { return E < 100.0 ? 0.0 : 1.0 + 0.25 * std::log(E / 100.0); }
double GNucShape(double E) {
  if (E <= 5.0) return 0.0;
  const double t = (E - 5.0) * (E - 5.0);
  const double onset = t / (t + 4.0);
  const double bump = 25.0 / ((E - 13.0) * (E - 13.0) + 25.0);
  return onset * (0.2 + bump);
}
}  // namespace

extern "C" int hepem_tables_set_nuclear(G4HepEmData* hd, double enuc_scale, double gnuc_scale) {
  if (!hd || !hd->fTheMaterialData || !hd->fTheElectronData || !hd->fThePositronData || !hd->fTheGammaData)
    return HEPEM_ERR_INVALID;
  if (!(enuc_scale >= 0.0) || !(gnuc_scale >= 0.0) || !std::isfinite(enuc_scale) || !std::isfinite(gnuc_scale))
    return HEPEM_ERR_INVALID;
  const G4HepEmMaterialData* md = hd->fTheMaterialData;
  const int nmat = md->fNumMaterialData;
  for (G4HepEmElectronData* ed : {hd->fTheElectronData, hd->fThePositronData}) {
    const int n = ed->fENucEnergyGridSize;
    if (ed->fNumMaterials != nmat || n < 2 || !ed->fENucEnergyGrid || !ed->fENucMacXsecData) return HEPEM_ERR_INVALID;
    std::vector<double> y(n), sd(n);
    for (int im = 0; im < nmat; ++im) {
      const double invX0 = 1.0 / md->fMaterialData[im].fRadiationLength;
      for (int i = 0; i < n; ++i) y[i] = enuc_scale * invX0 * ENucShape(ed->fENucEnergyGrid[i]);
      SplineSD(n, ed->fENucEnergyGrid, y.data(), sd.data());
      double* d = ed->fENucMacXsecData + (size_t)im * 2 * n;
      for (int i = 0; i < n; ++i) {
        d[2 * i] = y[i];
        d[2 * i + 1] = sd[i];
      }
    }
  }
  // window 2 stores [E, (total, total''), (conv, ''), (compton, ''), (PE, '')]
  // per node; the gamma-nuclear part is what the total holds beyond the sum
  // of the other three
  G4HepEmGammaData* gd = hd->fTheGammaData;
  if (gd->fNumMaterials != nmat || !gd->fMacXsecData || gd->fEGridSize2 < 2) return HEPEM_ERR_INVALID;
  const int n2 = gd->fEGridSize2;
  std::vector<double> x(n2), y(n2), sd(n2);
  for (int im = 0; im < nmat; ++im) {
    const double invX0 = 1.0 / md->fMaterialData[im].fRadiationLength;
    double* d = gd->fMacXsecData + (size_t)im * gd->fDataPerMat + gd->fNumData0 + gd->fNumData1;
    for (int i = 0; i < n2; ++i) {
      const double* r = d + 9 * i;
      x[i] = r[0];
      y[i] = r[3] + r[5] + r[7] + gnuc_scale * invX0 * GNucShape(r[0]);
    }
    SplineSD(n2, x.data(), y.data(), sd.data());
    for (int i = 0; i < n2; ++i) {
      d[9 * i + 1] = y[i];
      d[9 * i + 2] = sd[i];
    }
  }
  return HEPEM_OK;
}

extern "C" int hepem_tables_free(G4HepEmData* hd, G4HepEmParameters* pars) {
  if (pars) { std::free(pars->fParametersPerRegion); std::free(pars); }
  if (!hd) return HEPEM_OK;
  if (hd->fTheMatCutData) {
    std::free(hd->fTheMatCutData->fG4MCIndexToHepEmMCIndex);
    std::free(hd->fTheMatCutData->fMatCutData);
    std::free(hd->fTheMatCutData);
  }
  if (hd->fTheMaterialData) {
    for (int i = 0; i < hd->fTheMaterialData->fNumMaterialData; ++i) {
      G4HepEmMatData& d = hd->fTheMaterialData->fMaterialData[i];
      std::free(d.fElementVect); std::free(d.fNumOfAtomsPerVolumeVect);
      std::free(d.fSandiaEnergies); std::free(d.fSandiaCoefficients);
    }
    std::free(hd->fTheMaterialData->fG4MatIndexToHepEmMatIndex);
    std::free(hd->fTheMaterialData->fMaterialData);
    std::free(hd->fTheMaterialData);
  }
  if (hd->fTheElementData) {
    for (int z = 0; z <= hd->fTheElementData->fMaxZet; ++z) {
      std::free(hd->fTheElementData->fElementData[z].fSandiaEnergies);
      std::free(hd->fTheElementData->fElementData[z].fSandiaCoefficients);
    }
    std::free(hd->fTheElementData->fElementData);
    std::free(hd->fTheElementData);
  }
  for (G4HepEmElectronData* ed : {hd->fTheElectronData, hd->fThePositronData}) {
    if (!ed) continue;
    std::free(ed->fELossEnergyGrid); std::free(ed->fELossData);
    std::free(ed->fResMacXSecStartIndexPerMatCut); std::free(ed->fResMacXSecData);
    std::free(ed->fENucEnergyGrid); std::free(ed->fENucMacXsecData); std::free(ed->fTr1MacXSecData);
    std::free(ed->fElemSelectorIoniStartIndexPerMatCut); std::free(ed->fElemSelectorIoniData);
    std::free(ed->fElemSelectorBremSBStartIndexPerMatCut); std::free(ed->fElemSelectorBremSBData);
    std::free(ed->fElemSelectorBremRBStartIndexPerMatCut); std::free(ed->fElemSelectorBremRBData);
    std::free(ed);
  }
  if (hd->fTheSBTableData) {
    std::free(hd->fTheSBTableData->fGammaCutIndxStartIndexPerMC);
    std::free(hd->fTheSBTableData->fGammaCutIndices);
    std::free(hd->fTheSBTableData->fSBTableData);
    std::free(hd->fTheSBTableData);
  }
  if (hd->fTheGammaData) {
    G4HepEmGammaData* gd = hd->fTheGammaData;
    std::free(gd->fMacXsecData); std::free(gd->fElemSelectorConvStartIndexPerMat);
    std::free(gd->fElemSelectorConvEgrid); std::free(gd->fElemSelectorConvData);
    std::free(gd);
  }
  std::free(hd);
  return HEPEM_OK;
}

extern "C" int hepem_tables_serialize(const G4HepEmData* hd, const G4HepEmParameters* pars, uint8_t* buf, uint64_t* size) {
  if (!hd || !pars || !size) return HEPEM_ERR_INVALID;
  Writer w{buf};
  w.put(pars->fElectronTrackingCut); w.put(pars->fGammaTrackingCut); w.put(pars->fMinLossTableEnergy);
  w.put(pars->fMaxLossTableEnergy); w.put(pars->fNumLossTableBins); w.put(pars->fElectronBremModelLim);
  w.put(pars->fIsMSCPositronCor); w.put(pars->fIsMSCDisplacement);
  w.arr(pars->fParametersPerRegion, pars->fNumRegions);
  const G4HepEmMatCutData* mc = hd->fTheMatCutData;
  w.arr(mc->fG4MCIndexToHepEmMCIndex, mc->fNumG4MatCuts);
  w.arr(mc->fMatCutData, mc->fNumMatCutData);
  const G4HepEmMaterialData* md = hd->fTheMaterialData;
  for (int i = 0; i < md->fNumMaterialData; ++i) {
    const G4HepEmMatData& d = md->fMaterialData[i];
    w.arr(d.fElementVect, d.fNumOfElement); w.arr(d.fNumOfAtomsPerVolumeVect, d.fNumOfElement);
    w.put(d.fDensity); w.put(d.fDensityCorFactor); w.put(d.fElectronDensity); w.put(d.fRadiationLength);
    w.put(d.fMeanExEnergy); w.arr(d.fSandiaEnergies, d.fNumOfSandiaIntervals);
    w.arr(d.fSandiaCoefficients, 4 * d.fNumOfSandiaIntervals);
    w.put(d.fZeff); w.put(d.fZeff23); w.put(d.fZeffSqrt); w.put(d.fUMSCPar);
    w.arr(d.fUMSCStepMinPars, 2); w.arr(d.fUMSCTailCoeff, 4); w.arr(d.fUMSCThetaCoeff, 2);
  }
  const G4HepEmElementData* ed = hd->fTheElementData;
  for (int z = 0; z <= ed->fMaxZet; ++z) {
    const G4HepEmElemData& e = ed->fElementData[z];
    w.put(e.fZet);
    if (e.fZet < 0.0) continue;
    w.put(e.fZet13); w.put(e.fZet23); w.put(e.fCoulomb); w.put(e.fLogZ); w.put(e.fZFactor1);
    w.put(e.fDeltaMaxLow); w.put(e.fDeltaMaxHigh); w.put(e.fILVarS1); w.put(e.fILVarS1Cond);
    w.arr(e.fSandiaEnergies, e.fNumOfSandiaIntervals); w.arr(e.fSandiaCoefficients, 4 * e.fNumOfSandiaIntervals);
    w.put(e.fKShellBindingEnergy);
  }
  for (const G4HepEmElectronData* el : {hd->fTheElectronData, hd->fThePositronData}) {
    const int N = el->fELossEnergyGridSize;
    w.put(el->fELossLogMinEkin); w.put(el->fELossEILDelta);
    w.arr(el->fELossEnergyGrid, N); w.arr(el->fELossData, 5 * N * el->fNumMatCuts);
    w.arr(el->fResMacXSecStartIndexPerMatCut, el->fNumMatCuts); w.arr(el->fResMacXSecData, el->fResMacXSecNumData);
    w.put(el->fENucLogMinEkin); w.put(el->fENucEILDelta);
    w.arr(el->fENucEnergyGrid, el->fENucEnergyGridSize);
    w.arr(el->fENucMacXsecData, 2 * el->fENucEnergyGridSize * el->fNumMaterials);
    w.arr(el->fTr1MacXSecData, 2 * N * el->fNumMaterials);
    w.arr(el->fElemSelectorIoniStartIndexPerMatCut, el->fNumMatCuts); w.arr(el->fElemSelectorIoniData, el->fElemSelectorIoniNumData);
    w.arr(el->fElemSelectorBremSBStartIndexPerMatCut, el->fNumMatCuts); w.arr(el->fElemSelectorBremSBData, el->fElemSelectorBremSBNumData);
    w.arr(el->fElemSelectorBremRBStartIndexPerMatCut, el->fNumMatCuts); w.arr(el->fElemSelectorBremRBData, el->fElemSelectorBremRBNumData);
  }
  const G4HepEmSBTableData* sb = hd->fTheSBTableData;
  w.put(sb->fLogMinElEnergy); w.put(sb->fILDeltaElEnergy);
  w.arr(sb->fElEnergyVect, 65); w.arr(sb->fLElEnergyVect, 65); w.arr(sb->fKappaVect, 54); w.arr(sb->fLKappaVect, 54);
  w.arr(sb->fGammaCutIndxStartIndexPerMC, sb->fNumHepEmMatCuts); w.arr(sb->fGammaCutIndices, sb->fNumElemsInMatCuts);
  w.arr(sb->fSBTablesStartPerZ, 121); w.arr(sb->fSBTableData, sb->fNumSBTableData);
  const G4HepEmGammaData* gd = hd->fTheGammaData;
  w.put(gd->fEMin0); w.put(gd->fEMax0); w.put(gd->fLogEMin0); w.put(gd->fEILDelta0);
  w.put(gd->fEMax1); w.put(gd->fLogEMin1); w.put(gd->fEILDelta1);
  w.put(gd->fEMax2); w.put(gd->fLogEMin2); w.put(gd->fEILDelta2);
  w.arr(gd->fMacXsecData, (int64_t)gd->fNumMaterials * gd->fDataPerMat);
  w.put(gd->fElemSelectorConvLogMinEkin); w.put(gd->fElemSelectorConvEILDelta);
  w.arr(gd->fElemSelectorConvStartIndexPerMat, gd->fNumMaterials);
  w.arr(gd->fElemSelectorConvEgrid, gd->fElemSelectorConvEgridSize);
  w.arr(gd->fElemSelectorConvData, gd->fElemSelectorConvNumData);
  *size = w.pos;
  return HEPEM_OK;
}
