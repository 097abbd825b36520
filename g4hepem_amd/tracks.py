"""Structure-of-arrays track batches and step results, in host memory (numpy)
or HBM (torch tensors on the GPU), plus the synthetic workloads.

Arrays are planar: a field with k components holds k*n values, component c of
track i at [c*n + i], as the C-ABI expects.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._native import (HOWFAR_FIELDS, OUT_FIELDS, SUBSTEP_FIELDS, TRACK_FIELDS, HowFarState, StepOut, SubstepState,
                      Tracks, ELECTRON, POSITRON, GAMMA)


def _alloc(kind: str, dtype: str, count: int, device=None):
    if kind == "numpy":
        return np.zeros(count, dtype=dtype)
    import torch
    return torch.zeros(count, dtype=getattr(torch, dtype), device=device or "cuda")


def _ptr(a) -> int:
    return a.ctypes.data if isinstance(a, np.ndarray) else a.data_ptr()


class SoA:
    """A set of named planar arrays."""

    FIELDS: dict = {}

    def __init__(self, n: int, kind: str = "numpy", device=None, arrays: dict | None = None):
        self.n = int(n)
        self.kind = kind
        self.a = {}
        for name, (dt, k) in self.FIELDS.items():
            self.a[name] = arrays[name] if arrays and name in arrays else _alloc(kind, dt, k * self.n, device)

    def __getitem__(self, k):
        return self.a[k]

    def to_numpy(self) -> dict:
        return {k: (v if isinstance(v, np.ndarray) else v.cpu().numpy()) for k, v in self.a.items()}

    def nbytes(self) -> int:
        return sum(int(np.dtype(dt).itemsize) * k * self.n for dt, k in self.FIELDS.values())


class TrackSoA(SoA):
    FIELDS = TRACK_FIELDS

    def struct(self) -> Tracks:
        t = Tracks()
        t.n = self.n
        for name in self.FIELDS:
            setattr(t, name, _ptr(self.a[name]))
        return t

    @classmethod
    def from_numpy(cls, src: dict, kind: str = "numpy", device=None) -> "TrackSoA":
        n = len(src["ekin"])
        out = cls(n, kind, device)
        for k in cls.FIELDS:
            if kind == "numpy":
                out.a[k][...] = src[k]
            else:
                import torch
                out.a[k].copy_(torch.from_numpy(np.ascontiguousarray(src[k])))
        return out


class StepOutSoA(SoA):
    FIELDS = OUT_FIELDS

    def struct(self) -> StepOut:
        r = StepOut()
        for name in self.FIELDS:
            setattr(r, name, _ptr(self.a[name]))
        return r


class HowFarSoA(SoA):
    FIELDS = HOWFAR_FIELDS

    def struct(self) -> HowFarState:
        h = HowFarState()
        h.n = self.n
        for name in self.FIELDS:
            setattr(h, name, _ptr(self.a[name]))
        return h


class SubstepSoA(SoA):
    FIELDS = SUBSTEP_FIELDS

    def struct(self) -> SubstepState:
        h = SubstepState()
        h.n = self.n
        for name in self.FIELDS:
            setattr(h, name, _ptr(self.a[name]))
        return h


def step_outputs_differ(ptype, ta: TrackSoA, ra: StepOutSoA, tb: TrackSoA, rb: StepOutSoA) -> list:
    """Names of the outputs of a step -- results and updated track state -- in
    which two runs over the same batch differ bit for bit, compared on the
    device. `ptype` is the batch's particle-type tensor. Not compared, because a
    step leaves them unwritten: secondary slots at or beyond nsec, and the MSC
    state of gammas."""
    import torch
    n = ta.n
    bad = []
    nsec = ra["nsec"]
    el = ptype != GAMMA

    def bits(x):
        return x.view(torch.int64) if x.dtype == torch.float64 else x

    for name, (_, k) in OUT_FIELDS.items():
        a, b = bits(ra[name]), bits(rb[name])
        if name.startswith("sec_"):
            used = torch.stack([nsec > s for s in range(2)]).repeat(k // 2, 1).reshape(-1)
            a, b = a[used], b[used]
        if not torch.equal(a, b):
            bad.append(name)
    for name, (_, k) in TRACK_FIELDS.items():
        a, b = bits(ta[name]), bits(tb[name])
        if name.startswith("msc"):
            m = el.repeat(k)
            a, b = a[m], b[m]
        if not torch.equal(a, b):
            bad.append("state." + name)
    assert n == tb.n
    return bad


def reset_state_numpy(t: dict, n: int) -> None:
    """Start-of-track state (G4HepEmTrack/G4HepEmMSCTrackData ReSet, no cached Gaussian)."""
    t["num_ia_left"][:] = -1.0
    t["msc_init_range"][:] = 1.0e21
    t["msc_dyn_range_factor"][:] = 0.04
    t["msc_tlimit_min"][:] = 1.0e-7
    t["msc_first_step"][:] = 1
    t["rng_ctr"][:] = 0
    t["rng_gauss"][:] = 0.0
    t["rng_has_gauss"][:] = 0


# ---- synthetic workloads ------------------------------------------------------------------

def lookup_queries(n: int, seed: int, ncouples: int, lnemin=math.log(1e-3), lnemax=math.log(1e8), margin=0.0):
    """Config-1 style queries: couple ~ U{0..ncouples-1}, ln E ~ U[lnemin, lnemax]
    (optionally widened by `margin` on both sides, as the reference tests do to
    exercise the grid clamping); E and ln E are both passed."""
    rng = np.random.Generator(np.random.Philox(seed))
    imc = rng.integers(0, ncouples, n, dtype=np.int32)
    w = lnemax - lnemin
    lne = rng.uniform(lnemin - margin * w, lnemax + margin * w, n)
    return imc, np.exp(lne), lne


def step_tracks(n: int, seed: int, ncouples: int, lnemin=math.log(0.1), lnemax=math.log(1.0e4),
                mix=(4, 1, 5), safety_range=(1e-3, 10.0), first_id: int = 0) -> dict:
    """Config-2 style fresh tracks: particle type by deterministic interleave of
    e-:e+:gamma = mix per sum(mix) ids, ln E ~ U[lnemin, lnemax] MeV, isotropic
    direction, couple ~ U{0..ncouples-1}, safety ~ log-U[safety_range] mm, not on
    a boundary, start-of-track physics state."""
    rng = np.random.Generator(np.random.Philox(seed))
    ids = np.arange(first_id, first_id + n, dtype=np.int64)
    period = sum(mix)
    slot = ids % period
    ptype = np.where(slot < mix[0], ELECTRON, np.where(slot < mix[0] + mix[1], POSITRON, GAMMA)).astype(np.int8)
    t = {k: np.zeros(k_ * n, dtype=dt) for k, (dt, k_) in TRACK_FIELDS.items()}
    t["ptype"][:] = ptype
    t["id"][:] = ids.astype(np.int32)
    t["imc"][:] = rng.integers(0, ncouples, n, dtype=np.int32)
    t["ekin"][:] = np.exp(rng.uniform(lnemin, lnemax, n))
    cost = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * math.pi * rng.random(n)
    sint = np.sqrt((1.0 - cost) * (1.0 + cost))
    t["dir"][:] = np.concatenate([sint * np.cos(phi), sint * np.sin(phi), cost])
    t["safety"][:] = np.exp(rng.uniform(math.log(safety_range[0]), math.log(safety_range[1]), n))
    t["on_boundary"][:] = 0
    reset_state_numpy(t, n)
    return t


def type_ordered(t: dict) -> dict:
    """The same tracks with the e-/e+ first and the gammas after them, each in
    their original order: the layout a stack advance (hepem_stack_advance)
    hands to the next step."""
    n = len(t["ekin"])
    perm = np.argsort(t["ptype"] == GAMMA, kind="stable")
    return {k: np.ascontiguousarray(t[k].reshape(c, n)[:, perm]).reshape(-1) for k, (_, c) in TRACK_FIELDS.items()}
