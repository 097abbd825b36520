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
