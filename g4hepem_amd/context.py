"""Host-side handle of a device context: the Python mirror of the C-ABI.

Device memory and streams come from torch (plumbing only); every computation
runs in the HIP kernels of libg4hepem_amd.so.
"""
from __future__ import annotations

import ctypes as C

from . import _native as N
from .tracks import HowFarSoA, StepOutSoA, TrackSoA


def _stream_handle(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


class Context:
    """hepem_ctx: device image of the tables in FP64 or FP32-table form."""

    def __init__(self, tables: N.Tables, precision: int = N.HEPEM_FP64):
        self.tables = tables
        self.precision = precision
        h = C.c_void_p()
        N.check("hepem_ctx_create", N.lib().hepem_ctx_create(tables.data, tables.pars, precision, C.byref(h)))
        self.h = h

    def table_bytes(self) -> int:
        b = C.c_uint64()
        N.check("hepem_ctx_table_bytes", N.lib().hepem_ctx_table_bytes(self.h, C.byref(b)))
        return b.value

    def set_process_sort(self, on) -> None:
        """Sort of the pending interactions by process in the step and the
        Perform calls: True (always), False (never) or "auto" (the default:
        the faster choice as measured, which is always); results are
        identical."""
        mode = 2 if on == "auto" else (1 if on else 0)
        N.check("hepem_ctx_set_process_sort", N.lib().hepem_ctx_set_process_sort(self.h, mode))

    def set_lookup_staging(self, on: bool) -> None:
        """Config-1 lookups from tables staged in LDS (True, the default,
        where they fit in 64 kB) or from the table image; results are identical."""
        N.check("hepem_ctx_set_lookup_staging", N.lib().hepem_ctx_set_lookup_staging(self.h, 1 if on else 0))

    def lookup(self, ptype: int, imc, ekin, lekin, out=None, stream=None):
        """Config-1 lookups on device tensors; returns a [5, n] float64 tensor
        (dE/dx, range, inverse range, ioni and brem macroscopic cross sections)."""
        import torch
        n = imc.numel()
        if out is None:
            out = torch.empty(5 * n, dtype=torch.float64, device=imc.device)
        N.check("hepem_lookup_eloss_xsec",
                N.lib().hepem_lookup_eloss_xsec(self.h, ptype, n, imc.data_ptr(), ekin.data_ptr(), lekin.data_ptr(),
                                                out.data_ptr(), _stream_handle(stream)))
        return out.view(5, n)

    def lookup_gamma(self, imc, ekin, lekin, urnd, stream=None):
        """hepem_lookup_gamma_xsec on device tensors: returns (total
        macroscopic cross section, photoelectric cross section held by the
        track, selected interaction) tensors."""
        import torch
        n = imc.numel()
        xs = torch.empty(n, dtype=torch.float64, device=imc.device)
        pe = torch.empty(n, dtype=torch.float64, device=imc.device)
        pid = torch.empty(n, dtype=torch.int32, device=imc.device)
        N.check("hepem_lookup_gamma_xsec",
                N.lib().hepem_lookup_gamma_xsec(self.h, n, imc.data_ptr(), ekin.data_ptr(), lekin.data_ptr(),
                                                urnd.data_ptr(), xs.data_ptr(), pe.data_ptr(), pid.data_ptr(),
                                                _stream_handle(stream)))
        return xs, pe, pid

    def step(self, seed: int, tin: TrackSoA, tout: TrackSoA, res: StepOutSoA, work=None, stream=None) -> None:
        si, so, rr = tin.struct(), tout.struct(), res.struct()
        wp = work.data_ptr() if work is not None else None
        N.check("hepem_step", N.lib().hepem_step(self.h, C.c_uint64(seed), C.byref(si), C.byref(so), C.byref(rr), wp,
                                                  _stream_handle(stream)))

    def step_noalloc(self, seed: int, tin: TrackSoA, tout: TrackSoA, res: StepOutSoA, stream=None) -> None:
        """hepem_step_noalloc: the step with the library's fixed per-GPU work
        array (at most NOALLOC_MAX_TRACKS tracks; calls on one GPU must not
        overlap)."""
        si, so, rr = tin.struct(), tout.struct(), res.struct()
        N.check("hepem_step_noalloc", N.lib().hepem_step_noalloc(self.h, C.c_uint64(seed), C.byref(si), C.byref(so),
                                                                  C.byref(rr), _stream_handle(stream)))

    def step_prepare(self, tin: TrackSoA, work, stream=None) -> None:
        si = tin.struct()
        N.check("hepem_step_prepare", N.lib().hepem_step_prepare(self.h, C.byref(si), work.data_ptr(),
                                                                  _stream_handle(stream)))

    def step_run(self, seed: int, tin: TrackSoA, tout: TrackSoA, res: StepOutSoA, work, stream=None) -> None:
        si, so, rr = tin.struct(), tout.struct(), res.struct()
        N.check("hepem_step_run", N.lib().hepem_step_run(self.h, C.c_uint64(seed), C.byref(si), C.byref(so),
                                                          C.byref(rr), work.data_ptr(), _stream_handle(stream)))

    def howfar(self, seed: int, t: TrackSoA, hs: HowFarSoA, work=None, stream=None, kind: str = "mixed") -> None:
        """hepem_howfar (kind "mixed"), hepem_electron_howfar ("electron") or
        hepem_gamma_howfar ("gamma"): updates `t` in place and fills `hs`."""
        name = {"mixed": "hepem_howfar", "electron": "hepem_electron_howfar", "gamma": "hepem_gamma_howfar"}[kind]
        st, sh = t.struct(), hs.struct()
        wp = work.data_ptr() if work is not None else None
        N.check(name, getattr(N.lib(), name)(self.h, C.c_uint64(seed), C.byref(st), C.byref(sh), wp,
                                               _stream_handle(stream)))

    def perform(self, seed: int, t: TrackSoA, hs: HowFarSoA, res: StepOutSoA, work=None, stream=None,
                kind: str = "mixed") -> None:
        """hepem_perform / hepem_electron_perform / hepem_gamma_perform."""
        name = {"mixed": "hepem_perform", "electron": "hepem_electron_perform", "gamma": "hepem_gamma_perform"}[kind]
        st, sh, rr = t.struct(), hs.struct(), res.struct()
        wp = work.data_ptr() if work is not None else None
        N.check(name, getattr(N.lib(), name)(self.h, C.c_uint64(seed), C.byref(st), C.byref(sh), C.byref(rr), wp,
                                               _stream_handle(stream)))

    # TrackElectron's MSC sub-steps driven by the caller (hepem_electron_step_begin ...)
    def electron_step_begin(self, seed: int, t: TrackSoA, hs: HowFarSoA, ss, work, stream=None) -> None:
        self._substep("hepem_electron_step_begin", seed, t, hs, ss, work, stream)

    def electron_substep_howfar(self, seed: int, t: TrackSoA, hs: HowFarSoA, ss, work, stream=None) -> None:
        self._substep("hepem_electron_substep_howfar", seed, t, hs, ss, work, stream)

    def electron_substep_along(self, seed: int, t: TrackSoA, hs: HowFarSoA, ss, work, stream=None) -> None:
        self._substep("hepem_electron_substep_along", seed, t, hs, ss, work, stream)

    @staticmethod
    def electron_substep_active(work, stream=None) -> int:
        v = C.c_uint32()
        N.check("hepem_electron_substep_active",
                N.lib().hepem_electron_substep_active(work.data_ptr(), C.byref(v), _stream_handle(stream)))
        return v.value

    def electron_step_end(self, seed: int, t: TrackSoA, hs: HowFarSoA, ss, res: StepOutSoA, work, stream=None) -> None:
        st, sh, sss, rr = t.struct(), hs.struct(), ss.struct(), res.struct()
        N.check("hepem_electron_step_end",
                N.lib().hepem_electron_step_end(self.h, C.c_uint64(seed), C.byref(st), C.byref(sh), C.byref(sss),
                                                C.byref(rr), work.data_ptr(), _stream_handle(stream)))

    def _substep(self, name, seed, t, hs, ss, work, stream):
        st, sh, sss = t.struct(), hs.struct(), ss.struct()
        N.check(name, getattr(N.lib(), name)(self.h, C.c_uint64(seed), C.byref(st), C.byref(sh), C.byref(sss),
                                               work.data_ptr(), _stream_handle(stream)))

    def slab_run(self, setup: N.SlabSetup, nevents: int, seed: int, stream=None, per_event: bool = False) -> dict:
        """hepem_slab_run: showers in a stack of slabs. Returns per-slab sums
        over events of the deposit and its square, run totals and, with
        per_event, each event's deposits and leakage (1e-9 MeV units)."""
        import numpy as np
        nb = setup.nlayers * setup.nabs
        s1, s2 = np.zeros(nb), np.zeros(nb)
        r = N.SlabResult()
        r.edep_sum, r.edep_sum2 = s1.ctypes.data, s2.ctypes.data
        ev = lk = None
        if per_event:
            ev, lk = np.zeros(nevents * nb, dtype=np.int64), np.zeros(nevents, dtype=np.int64)
            r.edep_event, r.leak_event = ev.ctypes.data, lk.ctypes.data
        N.check("hepem_slab_run", N.lib().hepem_slab_run(self.h, C.byref(setup), nevents, C.c_uint64(seed), C.byref(r),
                                                         _stream_handle(stream)))
        out = {"edep_sum": s1, "edep_sum2": s2, "leak_sum": r.leak_sum, "edep_total": r.edep_total,
               "steps": r.steps, "tracks": r.tracks, "iterations": r.iterations, "max_in_flight": r.max_in_flight,
               "substeps": r.substeps}
        if per_event:
            out["edep_event"], out["leak_event"] = ev.reshape(nevents, nb), lk
        return out

    @staticmethod
    def rejected(work, stream=None) -> int:
        v = C.c_uint32()
        N.check("hepem_step_rejected", N.lib().hepem_step_rejected(work.data_ptr(), C.byref(v), _stream_handle(stream)))
        return v.value

    @staticmethod
    def work_buffer(n: int, device="cuda"):
        import torch
        return torch.empty(int(N.lib().hepem_step_work_bytes(n)), dtype=torch.uint8, device=device)

    def close(self):
        if getattr(self, "h", None):
            N.lib().hepem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan:
    """hepem_batch_plan: many track batches of uneven size stepped by one
    launch sequence (graph-capturable run)."""

    def __init__(self, ctx: Context, batches):
        """`batches`: sequence of (tin, tout, res) SoA triples on the device."""
        self.ctx = ctx
        self.batches = list(batches)  # keeps the arrays alive
        descs = (N.BatchDesc * len(self.batches))()
        for k, (tin, tout, res) in enumerate(self.batches):
            descs[k].inp = tin.struct()
            descs[k].out = tout.struct()
            descs[k].res = res.struct()
        h = C.c_void_p()
        N.check("hepem_batch_plan_create", N.lib().hepem_batch_plan_create(ctx.h, descs, len(self.batches), C.byref(h)))
        self.h = h

    def run(self, seed: int, stream=None) -> None:
        N.check("hepem_batch_plan_run", N.lib().hepem_batch_plan_run(self.h, C.c_uint64(seed), _stream_handle(stream)))

    def tune(self, seed: int, stream=None) -> tuple:
        N.check("hepem_batch_plan_tune", N.lib().hepem_batch_plan_tune(self.h, C.c_uint64(seed), _stream_handle(stream)))
        return self.shape()

    def shape(self) -> tuple:
        b, g = C.c_int32(), C.c_int32()
        N.check("hepem_batch_plan_shape", N.lib().hepem_batch_plan_shape(self.h, C.byref(b), C.byref(g)))
        return b.value, g.value

    def get_shape(self) -> dict:
        """The full launch shape: block, sort_tile (0: no process sort),
        persistent, grid."""
        x = N.BatchShape()
        N.check("hepem_batch_plan_get_shape", N.lib().hepem_batch_plan_get_shape(self.h, C.byref(x)))
        return {"block": x.block, "sort_tile": x.sort_tile, "persistent": x.persistent, "grid": x.grid}

    def set_shape(self, block: int, sort_tile: int, persistent: int) -> dict:
        x = N.BatchShape(block, sort_tile, persistent, 0)
        N.check("hepem_batch_plan_set_shape", N.lib().hepem_batch_plan_set_shape(self.h, C.byref(x)))
        return self.get_shape()

    @staticmethod
    def shapes(sort_modes=(0, 1)):
        """Every launch shape hepem_batch_plan_tune may choose from, as
        (block, sort_tile, persistent); sort_modes: 0 without, 1 with the
        process sort."""
        out = []
        for srt in sort_modes:
            for block in (64, 128, 256):
                for tile in ((1024, 2048, 4096, 8192) if srt else (0,)):
                    for pers in (0, 1):
                        out.append((block, tile, pers))
        return out

    def rejected(self, b: int, stream=None) -> int:
        v = C.c_uint32()
        N.check("hepem_batch_plan_rejected", N.lib().hepem_batch_plan_rejected(self.h, b, C.byref(v), _stream_handle(stream)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            N.lib().hepem_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _StackView:
    """The current batch of a TrackStack, in the stack's device memory: n and
    the hepem_tracks view (valid until the next push or advance)."""

    def __init__(self, t):
        self._t = t
        self.n = t.n

    def struct(self):
        return self._t


class TrackStack:
    """hepem_stack: a device track stack for callers that step batches with
    their own geometry (StackSecondaries for whole batches)."""

    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        h = C.c_void_p()
        N.check("hepem_stack_create", N.lib().hepem_stack_create(ctx.h, int(capacity), C.byref(h)))
        self.h = h

    def push(self, tracks: TrackSoA, stream=None) -> None:
        t = tracks.struct()
        N.check("hepem_stack_push", N.lib().hepem_stack_push(self.h, C.byref(t), _stream_handle(stream)))

    def view(self) -> _StackView:
        t = N.Tracks()
        N.check("hepem_stack_tracks", N.lib().hepem_stack_tracks(self.h, C.byref(t)))
        return _StackView(t)

    def __len__(self) -> int:
        return self.view().n

    def copy_out(self, dst: TrackSoA, stream=None) -> TrackSoA:
        t = dst.struct()
        N.check("hepem_stack_copy_out", N.lib().hepem_stack_copy_out(self.h, C.byref(t), _stream_handle(stream)))
        return dst

    def advance(self, out, res: StepOutSoA, keep=None, cut_edep=None, parent=None, origin=None, stream=None) -> int:
        """The next batch from a step of the current one (`out`: its
        post-step state, the view itself when stepped in place); keep,
        cut_edep, parent, origin: optional device arrays (uint8, float64,
        int32, int8). Returns the next batch's count."""
        t, r = out.struct(), res.struct()
        ptr = lambda a: a.data_ptr() if a is not None else None
        m = C.c_int32()
        N.check("hepem_stack_advance", N.lib().hepem_stack_advance(self.h, C.byref(t), C.byref(r), ptr(keep), ptr(cut_edep),
                                                                    ptr(parent), ptr(origin), C.byref(m),
                                                                    _stream_handle(stream)))
        return m.value

    def close(self):
        if getattr(self, "h", None):
            N.lib().hepem_stack_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
