"""The oracle's slab-shower harness (the CPU side of the config-5 comparison)
against what physics fixes: energy balance per event, photon transmission
through a thin slab (exp(-d * Sigma_tot) from the tables), and independence
from the number of host threads."""
import math

import numpy as np


def _setup(nlayers, thick, imc, primary, e0):
    from g4hepem_amd._native import slab_setup
    return slab_setup(nlayers, thick, imc, primary, e0, capacity=1 << 20, events_per_chunk=64)


def test_energy_balance_and_threads(tables1, oracle):
    su = _setup(50, (2.3, 5.7), (0, 1), 0, 300.0)
    d1, l1, s1, t1 = oracle.slab_shower(tables1, su, 24, 11, nthreads=1)
    d8, l8, s8, t8 = oracle.slab_shower(tables1, su, 24, 11, nthreads=8)
    assert np.array_equal(d1, d8) and np.array_equal(l1, l8) and (s1, t1) == (s8, t8)
    bal = (d1.sum(1) + l1) / 1e9
    assert np.abs(bal - 300.0).max() < 1e-5
    # a 300 MeV shower spreads over several radiation lengths and deposits in lAr too
    prof = d1.mean(0) / 1e9
    assert prof[1::2].sum() > 0.05 * prof.sum() and np.argmax(prof) > 2


def test_photon_transmission_closed_form(tables1, oracle):
    """1 MeV photons on one 2 mm Pb slab: an event deposits nothing exactly
    when the photon crosses without interacting (probability exp(-d Sigma))."""
    d, e0, nev = 2.0, 1.0, 20000
    su = _setup(1, (d,), (0,), 2, e0)
    dep, leak, _, _ = oracle.slab_shower(tables1, su, nev, 5, nthreads=8)
    zero = (dep.sum(1) == 0).mean()
    xs, _, _ = oracle.probe_gamma(tables1, np.array([0], np.int32), np.array([e0]), np.array([math.log(e0)]),
                                  np.array([0.5]))
    p = math.exp(-d * xs[0])
    sigma = math.sqrt(p * (1 - p) / nev)
    assert abs(zero - p) < 4 * sigma, (zero, p, sigma)
    # and those events leak exactly the primary energy
    assert np.all(leak[dep.sum(1) == 0] == int(round(e0 * 1e9)))


def _wdt_setup(nlayers, thick, imc, primary, e0, wmin):
    from g4hepem_amd._native import slab_setup
    return slab_setup(nlayers, thick, imc, primary, e0, capacity=1 << 20, events_per_chunk=64, woodcock_min_ekin=wmin)


def test_woodcock_transmission_two_materials(tables1, oracle):
    """Woodcock tracking (G4HepEmWoodcockHelper.cc:145-300) through 5.7 mm lAr
    followed by 2 mm Pb: the majorant is Pb's cross section, points in lAr are
    kept with probability Sigma_lAr / Sigma_Pb, and the gamma is taken out of
    the stack 1e-3 mm before its exit. The probability of crossing without
    interaction is exp(-(5.7 Sigma_lAr + (2 - 1e-3) Sigma_Pb))."""
    e0, nev = 1.0, 40000
    su = _wdt_setup(1, (5.7, 2.0), (1, 0), 2, e0, 0.2)
    dep, leak, steps, _ = oracle.slab_shower(tables1, su, nev, 6, nthreads=8)
    zero = (dep.sum(1) == 0).mean()
    xs, _, _ = oracle.probe_gamma(tables1, np.array([1, 0], np.int32), np.array([e0, e0]),
                                  np.array([math.log(e0)] * 2), np.array([0.5, 0.5]))
    p = math.exp(-(5.7 * xs[0] + (2.0 - 1.0e-3) * xs[1]))
    sigma = math.sqrt(p * (1 - p) / nev)
    assert abs(zero - p) < 4 * sigma, (zero, p, sigma)
    assert np.all(leak[dep.sum(1) == 0] == int(round(e0 * 1e9)))
    # a transmitted primary takes one step (no stop on the inner plane)
    su0 = _wdt_setup(1, (5.7, 2.0), (1, 0), 2, e0, 0.0)
    _, _, steps0, _ = oracle.slab_shower(tables1, su0, nev, 6, nthreads=8)
    assert steps < steps0


def test_woodcock_below_limit_is_normal_tracking(tables1, oracle):
    """Gammas at or below the Woodcock energy limit are tracked normally: a
    0.15 MeV photon shower with the limit at 0.2 MeV is the same, event by
    event, as with Woodcock tracking off."""
    a = oracle.slab_shower(tables1, _wdt_setup(10, (2.3, 5.7), (0, 1), 2, 0.15, 0.2), 2000, 3, nthreads=8)
    b = oracle.slab_shower(tables1, _wdt_setup(10, (2.3, 5.7), (0, 1), 2, 0.15, 0.0), 2000, 3, nthreads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2:] == b[2:]


def test_woodcock_shower_agrees_with_normal_tracking(tables1, oracle):
    """300 MeV e- showers in 50 x (2.3 mm Pb + 5.7 mm lAr): Woodcock tracking
    of the gammas above 0.2 MeV leaves the per-slab mean deposits unchanged
    (chi2 of the difference), conserves every event's energy and takes fewer
    steps."""
    nev, e0 = 600, 300.0
    runs = {}
    for w in (0.0, 0.2):
        dep, leak, steps, tracks = oracle.slab_shower(tables1, _wdt_setup(50, (2.3, 5.7), (0, 1), 0, e0, w), nev, 21,
                                                      nthreads=8)
        assert np.abs((dep.sum(1) + leak) / 1e9 - e0).max() < 1e-5
        runs[w] = (dep / 1e9, steps)
    a, b = runs[0.0][0], runs[0.2][0]
    v = a.var(0, ddof=1) / nev + b.var(0, ddof=1) / nev
    m = v > 0
    chi2, ndf = float((((a.mean(0) - b.mean(0)) ** 2)[m] / v[m]).sum()), int(m.sum())
    assert chi2 < ndf + 5 * math.sqrt(2 * ndf), (chi2, ndf)
    assert runs[0.2][1] < 0.8 * runs[0.0][1]
