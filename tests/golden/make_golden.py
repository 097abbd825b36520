"""Write the golden vectors of tests/golden/golden_v1.npz.

Runs the reference's own run-time code, compiled from the reference tree into
oracle/_ref/libhepem_ref.so (``make -C oracle ref``), on this project's tables
and the per-track Philox streams, and stores inputs and outputs:

* config-1 lookups (restricted dE/dx, range, inverse range, ioni/brem
  macroscopic cross sections) for e- and e+, 3 % outside the grid on both ends;
* single-function probes: stepping cross sections, transport mfp, nuclear and
  annihilation cross sections, the three element selectors, gamma total cross
  section + interaction sampling over all three energy windows;
* one HowFar+Perform step of fresh e-/e+/gamma tracks in three energy regimes;
* a 4-step chain (each step starting from the previous state) with tracks on
  and off volume boundaries;
* one step of tracks with 64-bit stream selectors (nonzero rng_stream);
* (golden_geom_v1.npz) a 3-step chain with a geometry between HowFar and
  Perform: each track's geometrical step is cut at a given distance, a cut
  step ends on a boundary and the next step starts from there;
* (golden_nuc_v1.npz) the 4-material set with non-zero e-/e+ nuclear and
  gamma-nuclear cross sections (Tables.set_nuclear(0.5, 0.2)): electron and
  gamma probes, one step in two energy regimes and a 3-step chain, in which
  e-/e+ winner 3 and gamma process 3 occur.

Usage:  python tests/golden/make_golden.py [all|nuc]
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from g4hepem_amd._native import Tables  # noqa: E402
from g4hepem_amd.tracks import lookup_queries, step_tracks  # noqa: E402
from oracle.oracle_py import Ref  # noqa: E402

OUT = Path(__file__).resolve().parent / "golden_v1.npz"
OUT_GEOM = Path(__file__).resolve().parent / "golden_geom_v1.npz"
OUT_NUC = Path(__file__).resolve().parent / "golden_nuc_v1.npz"
NUC_SCALES = (0.5, 0.2)
SET1 = ["G4_Pb", "G4_lAr"]
SET2 = ["G4_PbWO4", "G4_Pb", "G4_lAr", "G4_Si"]
CUT = 0.7
STEP_SEED = 0x5EED2026


def main():
    ref = Ref()
    t1 = Tables(SET1, CUT)
    t2 = Tables(SET2, CUT)
    g = {"digest_set1": np.array(t1.digest()), "digest_set2": np.array(t2.digest())}

    # config-1 lookups
    for pt, name in ((0, "e"), (1, "p")):
        imc, e, le = lookup_queries(2048, 100 + pt, len(SET1), margin=0.03)
        g[f"c1_{name}_imc"], g[f"c1_{name}_ekin"], g[f"c1_{name}_lekin"] = imc, e, le
        g[f"c1_{name}_out"] = ref.lookup(t1, pt, imc, e, le)

    # electron probes on the 4-material set (which: 0..5)
    for pt, name in ((0, "e"), (1, "p")):
        imc, e, le = lookup_queries(512, 200 + pt, len(SET2), lnemin=math.log(1e-4), lnemax=math.log(1e8), margin=0.03)
        g[f"pe_{name}_imc"], g[f"pe_{name}_ekin"], g[f"pe_{name}_lekin"] = imc, e, le
        g[f"pe_{name}_out"] = np.stack([ref.probe_electron(t2, pt, w, imc, e, le) for w in range(6)])

    # element selectors: PbWO4 (couple 0) is the multi-element material
    rng = np.random.Generator(np.random.Philox(300))
    ns = 1024
    for which, (lo, hi) in enumerate(((math.log(0.04), math.log(1e3)), (math.log(1e3), math.log(1e8)),
                                      (math.log(1.022), math.log(1e8)))):
        imc = np.zeros(ns, dtype=np.int32)
        le = rng.uniform(lo - 0.5, hi + 0.5, ns)
        u = rng.random(ns)
        g[f"sel{which}_lekin"], g[f"sel{which}_u"] = le, u
        g[f"sel{which}_out"] = ref.probe_selector(t2, 0, which, imc, np.exp(le), le, u)

    # gamma: total xsec + interaction sampling, all three windows
    imc, e, le = lookup_queries(2048, 400, len(SET2), lnemin=math.log(1e-4), lnemax=math.log(1e8), margin=0.01)
    u = np.random.Generator(np.random.Philox(401)).random(2048)
    xs, pe, pid = ref.probe_gamma(t2, imc, e, le, u)
    g.update(gm_imc=imc, gm_ekin=e, gm_lekin=le, gm_u=u, gm_xs=xs, gm_pe=pe, gm_pid=pid)

    # one step of fresh tracks in three regimes
    regimes = ((math.log(1e-3), math.log(1.0)), (math.log(0.1), math.log(1e4)), (math.log(1e3), math.log(1e8)))
    for r, (lo, hi) in enumerate(regimes):
        tr = step_tracks(768, 500 + r, len(SET2), lnemin=lo, lnemax=hi, first_id=10000 * r)
        to, res = ref.step(t2, STEP_SEED, tr, 8)
        for k, v in tr.items():
            g[f"st{r}_in_{k}"] = v
        for k, v in to.items():
            g[f"st{r}_out_{k}"] = v
        for k, v in res.items():
            g[f"st{r}_res_{k}"] = v

    # 4-step chain, 30 % of the tracks on a boundary
    tr = step_tracks(512, 600, len(SET2), lnemin=math.log(1e-2), lnemax=math.log(1e5), first_id=50000)
    tr["on_boundary"][:] = (np.random.Generator(np.random.Philox(601)).random(512) < 0.3).astype(np.uint8)
    for k, v in tr.items():
        g[f"ch_in_{k}"] = v
    cur = tr
    for s in range(4):
        cur, res = ref.step(t2, STEP_SEED + 1, cur, 8)
        for k, v in res.items():
            g[f"ch{s}_res_{k}"] = v
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
                  "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step"):
            g[f"ch{s}_out_{k}"] = cur[k]

    # one step with 64-bit stream selectors (rng_stream:id)
    tr = step_tracks(512, 650, len(SET2), lnemin=math.log(1e-2), lnemax=math.log(1e5), first_id=60000)
    tr["rng_stream"][:] = np.random.Generator(np.random.Philox(651)).integers(0, 2**32, 512, dtype=np.uint32)
    to, res = ref.step(t2, STEP_SEED + 3, tr, 8)
    for k, v in tr.items():
        g[f"hs_in_{k}"] = v
    for k, v in res.items():
        g[f"hs_res_{k}"] = v
    for k, v in to.items():
        g[f"hs_out_{k}"] = v

    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1024:.0f} KiB); set1 {t1.digest()[:16]} set2 {t2.digest()[:16]}")
    geometry_vectors(ref, t2)


def geometry_vectors(ref, t2, n=1024, nsteps=3):
    """Steps cut by a geometry: distances log-uniform in [1 um, 5 mm]."""
    g = {"digest_set2": np.array(t2.digest())}
    tr = step_tracks(n, 700, len(SET2), lnemin=math.log(1e-2), lnemax=math.log(1e4), first_id=70000)
    rng = np.random.Generator(np.random.Philox(701))
    for k, v in tr.items():
        g[f"gc_in_{k}"] = v
    cur = tr
    for s in range(nsteps):
        lim = np.exp(rng.uniform(math.log(1e-3), math.log(5.0), n))
        g[f"gc{s}_limit"] = lim
        cur, res = ref.step(t2, STEP_SEED + 2, cur, 8, geom_limit=lim)
        for k, v in res.items():
            g[f"gc{s}_res_{k}"] = v
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
                  "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step", "on_boundary"):
            g[f"gc{s}_out_{k}"] = cur[k]
    np.savez_compressed(OUT_GEOM, **g)
    print(f"wrote {OUT_GEOM} ({OUT_GEOM.stat().st_size / 1024:.0f} KiB)")


def nuclear_vectors(ref):
    """Tables with nuclear cross sections: probes, steps and a short chain."""
    t = Tables(SET2, CUT).set_nuclear(*NUC_SCALES)
    g = {"digest_nuc": np.array(t.digest())}
    for pt, name in ((0, "e"), (1, "p")):
        imc, e, le = lookup_queries(512, 800 + pt, len(SET2), lnemin=math.log(10.0), lnemax=math.log(1e8), margin=0.03)
        g[f"pe_{name}_imc"], g[f"pe_{name}_ekin"], g[f"pe_{name}_lekin"] = imc, e, le
        g[f"pe_{name}_out"] = np.stack([ref.probe_electron(t, pt, w, imc, e, le) for w in range(6)])
    imc, e, le = lookup_queries(2048, 810, len(SET2), lnemin=math.log(1e-4), lnemax=math.log(1e8), margin=0.01)
    u = np.random.Generator(np.random.Philox(811)).random(2048)
    xs, pe, pid = ref.probe_gamma(t, imc, e, le, u)
    g.update(gm_imc=imc, gm_ekin=e, gm_lekin=le, gm_u=u, gm_xs=xs, gm_pe=pe, gm_pid=pid)
    for r, (lo, hi) in enumerate(((math.log(0.1), math.log(1e4)), (math.log(1e2), math.log(1e6)))):
        tr = step_tracks(1024, 820 + r, len(SET2), lnemin=lo, lnemax=hi, first_id=80000 + 10000 * r)
        to, res = ref.step(t, STEP_SEED + 4, tr, 8)
        for k, v in tr.items():
            g[f"st{r}_in_{k}"] = v
        for k, v in to.items():
            g[f"st{r}_out_{k}"] = v
        for k, v in res.items():
            g[f"st{r}_res_{k}"] = v
    tr = step_tracks(512, 830, len(SET2), lnemin=math.log(1e2), lnemax=math.log(1e5), first_id=100000)
    tr["on_boundary"][:] = (np.random.Generator(np.random.Philox(831)).random(512) < 0.2).astype(np.uint8)
    for k, v in tr.items():
        g[f"ch_in_{k}"] = v
    cur = tr
    for s in range(3):
        cur, res = ref.step(t, STEP_SEED + 5, cur, 8)
        for k, v in res.items():
            g[f"ch{s}_res_{k}"] = v
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
                  "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step"):
            g[f"ch{s}_out_{k}"] = cur[k]
    np.savez_compressed(OUT_NUC, **g)
    print(f"wrote {OUT_NUC} ({OUT_NUC.stat().st_size / 1024:.0f} KiB); nuclear set {t.digest()[:16]}")


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "all":
        main()
    nuclear_vectors(Ref())
