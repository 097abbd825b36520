"""TrackElectron's MSC sub-step loop driven through the C-ABI
(hepem_electron_step_begin / _substep_howfar / _substep_along / _step_end) with
the caller's geometry written here in torch: each track at x in a slab
lo <= x <= hi across the x axis, crossing lo or hi leaving the world where its
exit bit says so, the MSC displacement applied off a boundary with the
reference's safety rule (G4HepEmTrackingManager.cc:523-566).

Against the oracle's restatement of the same loop (oracle_electron_slab_step,
itself pinned bit for bit to the loop over the reference's own compiled
functions): integers exact, floats within 1e-12 relative; with sub-steps off and
no way out of the world, against hepem_electron_howfar + hepem_electron_perform
with the same geometry, bit for bit.
"""
import math

import numpy as np
import pytest

from conftest import SET2, same_bits, valid_secondaries
from test_substep_cpu import slab_tracks, tables_multistep

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REL_TOL = 1e-12
ABS_TOL = 1e-12
K_MIN = 5.0e-8


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


class SlabGeometry:
    """The caller's geometry: a slab per track across x, on the device."""

    def __init__(self, x, lo, hi, exitf, dev):
        f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
        self.x, self.lo, self.hi = f(x), f(lo), f(hi)
        ex = torch.from_numpy(np.ascontiguousarray(exitf, np.uint8)).to(dev)
        self.exit_lo, self.exit_hi = (ex & 1) != 0, (ex & 2) != 0
        self.side = torch.zeros(len(x), dtype=torch.int8, device=dev)

    def safety(self):
        return torch.clamp(torch.minimum(self.x - self.lo, self.hi - self.x), min=0.0)

    def move(self, t, hs, ss, n):
        """Moves the active tracks at most hs.geom_step along x; writes the
        length moved, the post-move boundary flag and left_world."""
        act = ss["active"].bool()
        ux = t["dir"][:n]
        gs = hs["geom_step"]
        inf = torch.full_like(gs, math.inf)
        # + 0.0: a track on the plane it moves out through gets +0 (as a navigator's distance)
        dist = torch.where(ux > 0, (self.hi - self.x) / ux, torch.where(ux < 0, (self.lo - self.x) / ux, inf)) + 0.0
        limited = act & (dist < gs)
        final = torch.where(limited, dist, gs)
        plane = torch.where(ux > 0, self.hi, self.lo)
        moved = torch.where(limited, plane, torch.clamp(self.x + final * ux, self.lo, self.hi))
        self.x = torch.where(act, moved, self.x)
        hs["geom_step"].copy_(torch.where(act, final, gs))
        t["on_boundary"].copy_(torch.where(act, limited, t["on_boundary"].bool()).to(torch.uint8))
        sd = torch.where(ux > 0, 1, -1).to(torch.int8)
        self.side = torch.where(limited, sd, self.side)
        left = limited & (((ux > 0) & self.exit_hi) | ((ux < 0) & self.exit_lo))
        ss["left_world"].copy_(torch.where(act, left, ss["left_world"].bool()).to(torch.uint8))
        return act, limited

    def displace(self, ss, took, limited, n):
        d = ss["displacement"]
        d0, d1, d2 = d[:n], d[n:2 * n], d[2 * n:]
        dl2 = d0 * d0 + d1 * d1 + d2 * d2
        disp_r = torch.sqrt(dl2)
        ps = 0.99 * self.safety()
        nx = torch.where((ps > 0) & (disp_r <= ps), self.x + d0,
                         torch.where(disp_r < ps, self.x + d0,
                                     torch.where(ps > K_MIN, self.x + d0 * (ps / disp_r), self.x)))
        self.x = torch.where(took & ~limited & (dl2 > K_MIN * K_MIN), nx, self.x)


def run_substeps(ctx, seed, tr, x, lo, hi, exitf, dev):
    from g4hepem_amd.tracks import HowFarSoA, StepOutSoA, SubstepSoA, TrackSoA
    n = len(tr["ekin"])
    t = TrackSoA.from_numpy(tr, "torch", dev)
    geo = SlabGeometry(x, lo, hi, exitf, dev)
    t["safety"].copy_(torch.where(t["on_boundary"].bool(), torch.zeros_like(t["safety"]), geo.safety()))
    hs, ss, res = HowFarSoA(n, "torch", dev), SubstepSoA(n, "torch", dev), StepOutSoA(n, "torch", dev)
    work = ctx.work_buffer(0, dev)
    ctx.electron_step_begin(seed, t, hs, ss, work)
    active = ctx.electron_substep_active(work)
    rounds = 0
    while active > 0:
        ctx.electron_substep_howfar(seed, t, hs, ss, work)
        took, limited = geo.move(t, hs, ss, n)
        ctx.electron_substep_along(seed, t, hs, ss, work)
        geo.displace(ss, took, limited, n)
        active = ctx.electron_substep_active(work)
        rounds += 1
        assert rounds < 100_000
    ctx.electron_step_end(seed, t, hs, ss, res, work)
    torch.cuda.synchronize()
    g = {"x": geo.x.cpu().numpy(), "side": geo.side.cpu().numpy(), "left": ss["left_world"].cpu().numpy(),
         "nsub": ss["nsub"].cpu().numpy()}
    return t.to_numpy(), valid_secondaries(res.to_numpy()), g, rounds


def close(a, b, absolute=0.0):
    return np.all(np.abs(a - b) <= REL_TOL * np.abs(b) + absolute)


@pytest.mark.parametrize("multi", [True, False], ids=["substeps", "one-step"])
def test_substep_calls_match_oracle(tables2, oracle, dev, multi):
    from g4hepem_amd import HEPEM_FP64, Context
    tables = tables2 if multi else tables_multistep(tables2, False)
    tr, x, lo, hi, ex = slab_tracks(200_000, 5, len(SET2))
    _print(f"substep calls, multi={multi}: gpu")
    g_t, g_r, g_g, rounds = run_substeps(Context(tables, HEPEM_FP64), 77, tr, x, lo, hi, ex, dev)
    _print(f"  {rounds} sub-step rounds; oracle")
    o_t, o_r, o_g = oracle.electron_slab_step(tables, 77, tr, x, lo, hi, ex, nthreads=16)
    o_r = valid_secondaries(o_r)
    for k in ("winner", "stopped", "nsec", "sec_ptype"):
        assert same_bits(g_r[k], o_r[k]), k
    for k in ("side", "left", "nsub"):
        assert same_bits(g_g[k], o_g[k]), k
    for k in ("rng_ctr", "rng_has_gauss", "msc_first_step"):
        assert same_bits(g_t[k], o_t[k]), k
    # a track that left the world keeps no boundary flag of its own (TrackElectron
    # kills it before setting one): compare the others
    stay = o_g["left"] == 0
    assert same_bits(g_t["on_boundary"][stay], o_t["on_boundary"][stay])
    for k in ("true_step", "geom_step", "edep", "sec_ekin"):
        assert close(g_r[k], o_r[k]), k
    assert close(g_r["sec_dir"], o_r["sec_dir"], ABS_TOL)
    for k in ("ekin", "num_ia_left", "msc_init_range", "msc_dyn_range_factor", "msc_tlimit_min", "rng_gauss"):
        assert close(g_t[k], o_t[k]), k
    assert close(g_t["dir"], o_t["dir"], ABS_TOL)
    assert close(g_g["x"], o_g["x"], ABS_TOL)
    exact = np.mean([same_bits(g_r[k], o_r[k]) for k in ("edep", "true_step")])
    _print(f"  nsub max {g_g['nsub'].max()}, left {g_g['left'].mean():.3f}, edep/true_step bit-exact: {exact}")
    if multi:
        assert (g_g["nsub"] > 1).mean() > 0.05
    else:
        assert (g_g["nsub"] == 1).all()


def test_substep_calls_without_substeps_equal_howfar_perform(tables2, dev):
    """Sub-steps off, no way out of the world: begin + one sub-step + end is
    HowFar, the geometry and Perform, bit for bit."""
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd.tracks import HowFarSoA, StepOutSoA, TrackSoA
    tables = tables_multistep(tables2, False)
    ctx = Context(tables, HEPEM_FP64)
    tr, x, lo, hi, ex = slab_tracks(100_000, 13, len(SET2), exits=False)
    a_t, a_r, a_g, _ = run_substeps(ctx, 9, tr, x, lo, hi, ex, dev)
    n = len(x)
    t = TrackSoA.from_numpy(tr, "torch", dev)
    geo = SlabGeometry(x, lo, hi, ex, dev)
    t["safety"].copy_(torch.where(t["on_boundary"].bool(), torch.zeros_like(t["safety"]), geo.safety()))
    hs, res = HowFarSoA(n, "torch", dev), StepOutSoA(n, "torch", dev)
    work = ctx.work_buffer(n, dev)
    ctx.howfar(9, t, hs, work, kind="electron")
    ux = t["dir"][:n]
    gs = hs["geom_step"]
    dist = torch.where(ux > 0, (geo.hi - geo.x) / ux, torch.where(ux < 0, (geo.lo - geo.x) / ux, torch.full_like(gs, math.inf))) + 0.0
    limited = dist < gs
    hs["geom_step"].copy_(torch.where(limited, dist, gs))
    t["on_boundary"].copy_(limited.to(torch.uint8))
    ctx.perform(9, t, hs, res, work, kind="electron")
    torch.cuda.synchronize()
    b_t, b_r = t.to_numpy(), valid_secondaries(res.to_numpy())
    for k in b_r:
        if k != "displacement":   # applied by the caller after the sub-step calls
            assert same_bits(a_r[k], b_r[k]), k
    for k in b_t:
        assert same_bits(a_t[k], b_t[k]), k


def test_substep_calls_reject_other_slots(tables2, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd.tracks import step_tracks
    tr = step_tracks(1000, 3, len(SET2))          # e-, e+ and gammas
    n = len(tr["ekin"])
    x, lo, hi, ex = np.full(n, 0.5), np.zeros(n), np.ones(n), np.zeros(n, np.uint8)
    g_t, g_r, g_g, _ = run_substeps(Context(tables2, HEPEM_FP64), 1, tr, x, lo, hi, ex, dev)
    gam = tr["ptype"] == 2
    assert (g_r["winner"][gam] == -99).all() and (g_r["winner"][~gam] != -99).all()
    assert same_bits(g_t["ekin"][gam], tr["ekin"][gam])
