"""Gamma macroscopic cross sections and interaction selection on the GPU
(hepem_lookup_gamma_xsec, the batched TestMacXSecDataKernel of
testing/GammaXSections/src/GammaMacXSecs.cu:20-31) through the C-ABI.

FP64 form: bit-exact against the golden vectors the compiled reference wrote
and against the oracle over 1 keV - 100 TeV, on the plain tables and on the
tables with a gamma-nuclear cross section (process 3 selected).
FP32-table form: total and photoelectric cross sections within 1e-5 relative
of the FP64 oracle in each of the three energy windows of GetTotalMacXSec
(G4HepEmGammaManager.icc:108-153: E <= 150 keV Compton + Sandia PE,
150 keV - 2mc^2 the linear 2-column table, above 2mc^2 the 4-column spline).
"""
import math

import numpy as np
import pytest

from conftest import CUT, SET2, same_bits

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FP32_REL_TOL = 1e-5
E_WINDOW0 = 0.15                  # MeV: upper end of the Compton + Sandia PE window
E_WINDOW1 = 2.0 * 0.51099895      # MeV: 2 mc^2, upper end of the linear window


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


def gpu_gamma(ctx, imc, ekin, lekin, u, dev):
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (imc.astype(np.int32), ekin, lekin, u)]
    xs, pe, pid = ctx.lookup_gamma(*t)
    torch.cuda.synchronize()
    return xs.cpu().numpy(), pe.cpu().numpy(), pid.cpu().numpy()


def queries(n, seed, ncouples, lo=1e-3, hi=1e8):
    rng = np.random.Generator(np.random.Philox(seed))
    le = rng.uniform(math.log(lo), math.log(hi), n)
    return rng.integers(0, ncouples, n).astype(np.int32), np.exp(le), le, rng.random(n)


def test_gamma_golden_bitexact(golden, tables2, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    ctx = Context(tables2, HEPEM_FP64)
    xs, pe, pid = gpu_gamma(ctx, golden["gm_imc"], golden["gm_ekin"], golden["gm_lekin"], golden["gm_u"], dev)
    assert same_bits(xs, golden["gm_xs"])
    assert same_bits(pid, golden["gm_pid"])
    known = pid != 3
    assert same_bits(pe[known], golden["gm_pe"][known])


@pytest.mark.parametrize("nuclear", [False, True], ids=["plain", "gamma-nuclear"])
def test_gamma_fp64_equals_oracle(tables2, oracle, dev, nuclear):
    from g4hepem_amd import HEPEM_FP64, Context, Tables
    tables = Tables(SET2, CUT).set_nuclear(0.5, 0.2) if nuclear else tables2
    ctx = Context(tables, HEPEM_FP64)
    imc, e, le, u = queries(200_000, 31, len(SET2))
    imc[:3] = (-1, 4, 99)   # invalid couples: NaN and -99, no table read
    xs, pe, pid = gpu_gamma(ctx, imc, e, le, u, dev)
    assert np.isnan(xs[:3]).all() and (pid[:3] == -99).all()
    rxs, rpe, rpid = oracle.probe_gamma(tables, imc[3:], e[3:], le[3:], u[3:])
    assert same_bits(xs[3:], rxs) and same_bits(pid[3:], rpid) and same_bits(pe[3:], rpe)
    expect = {0, 1, 2, 3} if nuclear else {0, 1, 2}
    assert set(np.unique(pid[3:]).tolist()) == expect


def test_gamma_fp32_errors_per_window(tables2, oracle, dev):
    from g4hepem_amd import HEPEM_FP32T, Context
    ctx = Context(tables2, HEPEM_FP32T)
    imc, e, le, u = queries(400_000, 32, len(SET2))
    xs, pe, pid = gpu_gamma(ctx, imc, e, le, u, dev)
    rxs, rpe, rpid = oracle.probe_gamma(tables2, imc, e, le, u)
    windows = {"compton+sandia": e <= E_WINDOW0, "linear": (e > E_WINDOW0) & (e <= E_WINDOW1), "spline": e > E_WINDOW1}
    for name, w in windows.items():
        assert w.sum() > 10_000, name
        m = w & (rxs > 0)
        rel = np.abs(xs[m] - rxs[m]) / rxs[m]
        _print(f"{name}: {m.sum()} queries, total xsec max rel {rel.max():.3e} mean {rel.mean():.3e}")
        assert rel.max() <= FP32_REL_TOL, name
        if name != "spline":   # below 2mc^2 the track's PE value is the PE cross section itself
            mp = w & (rpe > 0)
            relp = np.abs(pe[mp] - rpe[mp]) / rpe[mp]
            _print(f"{name}: PE xsec max rel {relp.max():.3e}")
            assert relp.max() <= FP32_REL_TOL, name
    agree = (pid == rpid).mean()
    _print(f"selected interaction agreement {agree:.6f}")
    assert agree > 0.9999
