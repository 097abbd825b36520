"""ctypes binding of the C-ABI in include/g4hepem_amd.h and
include/g4hepem_amd_tables.h.

The library is built in-tree (``python -m g4hepem_amd.build``); importing this
module without it raises, so no code path silently falls back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libg4hepem_amd.so"

HEPEM_FP64 = 0
HEPEM_FP32T = 1
ELECTRON, POSITRON, GAMMA = 0, 1, 2

NOALLOC_MAX_TRACKS = 262144   # HEPEM_NOALLOC_MAX_TRACKS

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "capacity exceeded",
          -4: "allocation failed", -5: "no usable GPU device", -6: "unsupported option"}


class HepEmError(RuntimeError):
    def __init__(self, where: str, status: int):
        super().__init__(f"{where} failed: {STATUS.get(status, status)} ({status})")
        self.status = status


P = C.c_void_p


class Tracks(C.Structure):
    """hepem_tracks: per-track state (device or host pointers)."""
    _fields_ = [("n", C.c_int32), ("ptype", P), ("id", P), ("imc", P), ("ekin", P), ("dir", P),
                ("safety", P), ("on_boundary", P), ("num_ia_left", P), ("msc_init_range", P),
                ("msc_dyn_range_factor", P), ("msc_tlimit_min", P), ("msc_first_step", P),
                ("rng_ctr", P), ("rng_gauss", P), ("rng_has_gauss", P), ("rng_stream", P)]


class StepOut(C.Structure):
    """hepem_step_out: per-step results."""
    _fields_ = [("true_step", P), ("geom_step", P), ("edep", P), ("displacement", P), ("winner", P),
                ("stopped", P), ("nsec", P), ("sec_ptype", P), ("sec_ekin", P), ("sec_dir", P)]


class HowFarState(C.Structure):
    """hepem_howfar_state: HowFar -> Perform hand-over of one step."""
    _fields_ = [("n", C.c_int32), ("geom_step", P), ("range", P), ("mfp", P), ("winner", P),
                ("msc_lambtr1", P), ("msc_true_step", P), ("msc_zpath", P), ("msc_par1", P),
                ("msc_par2", P), ("msc_par3", P), ("msc_active", P), ("msc_displace", P), ("pe_xsec", P)]


class SlabSetup(C.Structure):
    """hepem_slab_setup (include/g4hepem_amd_shower.h)."""
    _fields_ = [("nlayers", C.c_int32), ("nabs", C.c_int32), ("thickness", C.c_double * 4), ("imc", C.c_int32 * 4),
                ("primary", C.c_int32), ("ekin0", C.c_double), ("capacity", C.c_int32), ("events_per_chunk", C.c_int32),
                ("refill_below", C.c_int32), ("woodcock_min_ekin", C.c_double)]


class SlabResult(C.Structure):
    _fields_ = [("edep_sum", P), ("edep_sum2", P), ("edep_event", P), ("leak_event", P), ("leak_sum", C.c_double), ("edep_total", C.c_double),
                ("steps", C.c_int64), ("tracks", C.c_int64), ("iterations", C.c_int32), ("max_in_flight", C.c_int32),
                ("substeps", C.c_int64)]


def slab_setup(nlayers: int, thickness, imc, primary: int, ekin0: float, capacity: int = 1 << 22,
               events_per_chunk: int = 256, refill_below: int = 0, woodcock_min_ekin: float = 0.0) -> SlabSetup:
    s = SlabSetup()
    s.nlayers, s.nabs = nlayers, len(thickness)
    for k, (t, m) in enumerate(zip(thickness, imc)):
        s.thickness[k], s.imc[k] = t, m
    s.primary, s.ekin0, s.capacity, s.events_per_chunk = primary, ekin0, capacity, events_per_chunk
    s.refill_below = refill_below
    s.woodcock_min_ekin = woodcock_min_ekin
    return s


class SubstepState(C.Structure):
    """hepem_substep_state: TrackElectron's loop state of a batch of e-/e+."""
    _fields_ = [("n", C.c_int32), ("discrete_winner", P), ("step_left", P), ("true_path", P), ("geom_path", P),
                ("mean_eloss", P), ("pre_ekin", P), ("displacement", P), ("active", P), ("more", P), ("stopped", P),
                ("left_world", P), ("nsub", P)]


class BatchDesc(C.Structure):
    _fields_ = [("inp", Tracks), ("out", Tracks), ("res", StepOut)]


class BatchShape(C.Structure):
    """hepem_batch_shape: launch shape of a batch plan's step."""
    _fields_ = [("block", C.c_int32), ("sort_tile", C.c_int32), ("persistent", C.c_int32), ("grid", C.c_int32)]


# field name -> (dtype name, components)
TRACK_FIELDS = {
    "ptype": ("int8", 1), "id": ("int32", 1), "imc": ("int32", 1), "ekin": ("float64", 1),
    "dir": ("float64", 3), "safety": ("float64", 1), "on_boundary": ("uint8", 1),
    "num_ia_left": ("float64", 4), "msc_init_range": ("float64", 1),
    "msc_dyn_range_factor": ("float64", 1), "msc_tlimit_min": ("float64", 1),
    "msc_first_step": ("uint8", 1), "rng_ctr": ("uint32", 1), "rng_gauss": ("float64", 1),
    "rng_has_gauss": ("uint8", 1), "rng_stream": ("uint32", 1),
}
OUT_FIELDS = {
    "true_step": ("float64", 1), "geom_step": ("float64", 1), "edep": ("float64", 1),
    "displacement": ("float64", 3), "winner": ("int32", 1), "stopped": ("uint8", 1),
    "nsec": ("int32", 1), "sec_ptype": ("int8", 2), "sec_ekin": ("float64", 2), "sec_dir": ("float64", 6),
}

SUBSTEP_FIELDS = {
    "discrete_winner": ("int32", 1), "step_left": ("float64", 1), "true_path": ("float64", 1),
    "geom_path": ("float64", 1), "mean_eloss": ("float64", 1), "pre_ekin": ("float64", 1),
    "displacement": ("float64", 3), "active": ("uint8", 1), "more": ("uint8", 1), "stopped": ("uint8", 1),
    "left_world": ("uint8", 1), "nsub": ("int32", 1),
}

HOWFAR_FIELDS = {
    "geom_step": ("float64", 1), "range": ("float64", 1), "mfp": ("float64", 4), "winner": ("int32", 1),
    "msc_lambtr1": ("float64", 1), "msc_true_step": ("float64", 1), "msc_zpath": ("float64", 1),
    "msc_par1": ("float64", 1), "msc_par2": ("float64", 1), "msc_par3": ("float64", 1),
    "msc_active": ("uint8", 1), "msc_displace": ("uint8", 1), "pe_xsec": ("float64", 1),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m g4hepem_amd.build native`")
    L = C.CDLL(str(LIB_PATH))
    i32, u64, vp, dp, ip = C.c_int32, C.c_uint64, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.hepem_tables_build.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.c_double, C.POINTER(vp), C.POINTER(vp)]
    L.hepem_tables_free.argtypes = [vp, vp]
    L.hepem_tables_set_nuclear.argtypes = [vp, C.c_double, C.c_double]
    L.hepem_tables_serialize.argtypes = [vp, vp, vp, C.POINTER(u64)]
    L.hepem_tables_from_json.argtypes = [C.c_char_p, u64, C.POINTER(vp), C.POINTER(vp)]
    L.hepem_tables_to_json.argtypes = [vp, vp, vp, C.POINTER(u64)]
    L.hepem_ctx_create.argtypes = [vp, vp, C.c_int, C.POINTER(vp)]
    L.hepem_ctx_destroy.argtypes = [vp]
    L.hepem_ctx_table_bytes.argtypes = [vp, C.POINTER(u64)]
    L.hepem_status_string.restype = C.c_char_p
    L.hepem_lookup_eloss_xsec.argtypes = [vp, C.c_int, i32, vp, vp, vp, vp, vp]
    L.hepem_step.argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(Tracks), C.POINTER(StepOut), vp, vp]
    L.hepem_lookup_gamma_xsec.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.hepem_step_noalloc.argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(Tracks), C.POINTER(StepOut), vp]
    L.hepem_step_work_bytes.argtypes = [i32]
    L.hepem_step_work_bytes.restype = u64
    L.hepem_step_prepare.argtypes = [vp, C.POINTER(Tracks), vp, vp]
    L.hepem_step_run.argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(Tracks), C.POINTER(StepOut), vp, vp]
    L.hepem_step_rejected.argtypes = [vp, C.POINTER(C.c_uint32), vp]
    L.hepem_ctx_set_process_sort.argtypes = [vp, C.c_int]
    L.hepem_ctx_set_lookup_staging.argtypes = [vp, C.c_int]
    L.hepem_step_batched.argtypes = [vp, u64, C.POINTER(BatchDesc), i32, vp]
    L.hepem_slab_run.argtypes = [vp, C.POINTER(SlabSetup), i32, u64, C.POINTER(SlabResult), vp]
    L.hepem_batch_plan_create.argtypes = [vp, C.POINTER(BatchDesc), i32, C.POINTER(vp)]
    L.hepem_batch_plan_destroy.argtypes = [vp]
    L.hepem_batch_plan_run.argtypes = [vp, u64, vp]
    L.hepem_batch_plan_tune.argtypes = [vp, u64, vp]
    L.hepem_batch_plan_shape.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    L.hepem_batch_plan_get_shape.argtypes = [vp, C.POINTER(BatchShape)]
    L.hepem_stack_create.argtypes = [vp, i32, C.POINTER(vp)]
    L.hepem_stack_destroy.argtypes = [vp]
    L.hepem_stack_tracks.argtypes = [vp, C.POINTER(Tracks)]
    L.hepem_stack_push.argtypes = [vp, C.POINTER(Tracks), vp]
    L.hepem_stack_copy_out.argtypes = [vp, C.POINTER(Tracks), vp]
    L.hepem_stack_advance.argtypes = [vp, C.POINTER(Tracks), C.POINTER(StepOut), vp, vp, vp, vp, C.POINTER(i32), vp]
    L.hepem_batch_plan_set_shape.argtypes = [vp, C.POINTER(BatchShape)]
    L.hepem_batch_plan_rejected.argtypes = [vp, i32, C.POINTER(C.c_uint32), vp]
    L.hepem_tracks_reset_state.argtypes = [C.POINTER(Tracks), vp]
    for name in ("hepem_electron_step_begin", "hepem_electron_substep_howfar", "hepem_electron_substep_along"):
        getattr(L, name).argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(HowFarState), C.POINTER(SubstepState), vp, vp]
    L.hepem_electron_step_end.argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(HowFarState), C.POINTER(SubstepState),
                                          C.POINTER(StepOut), vp, vp]
    L.hepem_electron_substep_active.argtypes = [vp, C.POINTER(C.c_uint32), vp]
    split = ("hepem_howfar", "hepem_electron_howfar", "hepem_gamma_howfar")
    for name in split:
        getattr(L, name).argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(HowFarState), vp, vp]
    for name in ("hepem_perform", "hepem_electron_perform", "hepem_gamma_perform"):
        getattr(L, name).argtypes = [vp, u64, C.POINTER(Tracks), C.POINTER(HowFarState), C.POINTER(StepOut), vp, vp]
    for name in ("hepem_tables_build", "hepem_tables_free", "hepem_tables_set_nuclear", "hepem_tables_serialize", "hepem_tables_from_json",
                 "hepem_tables_to_json", "hepem_ctx_create",
                 "hepem_ctx_destroy", "hepem_ctx_table_bytes", "hepem_lookup_eloss_xsec", "hepem_lookup_gamma_xsec", "hepem_step", "hepem_step_noalloc",
                 "hepem_step_prepare", "hepem_step_run", "hepem_step_rejected", "hepem_ctx_set_process_sort",
                 "hepem_ctx_set_lookup_staging",
                 "hepem_step_batched", "hepem_tracks_reset_state", "hepem_howfar", "hepem_perform",
                 "hepem_electron_howfar", "hepem_electron_perform", "hepem_gamma_howfar", "hepem_gamma_perform",
                 "hepem_batch_plan_create", "hepem_batch_plan_destroy", "hepem_batch_plan_run", "hepem_batch_plan_tune",
                 "hepem_batch_plan_shape", "hepem_batch_plan_get_shape", "hepem_batch_plan_set_shape",
                 "hepem_stack_create", "hepem_stack_destroy", "hepem_stack_tracks", "hepem_stack_push",
                 "hepem_stack_copy_out", "hepem_stack_advance",
                 "hepem_batch_plan_rejected", "hepem_slab_run", "hepem_electron_step_begin",
                 "hepem_electron_substep_howfar", "hepem_electron_substep_along", "hepem_electron_substep_active",
                 "hepem_electron_step_end"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(where: str, rc: int) -> None:
    if rc != 0:
        raise HepEmError(where, rc)


class Tables:
    """G4HepEmData + G4HepEmParameters: built by the offline table builder, or
    read from a G4HepEmState JSON document (Tables.from_json)."""

    def __init__(self, materials, prod_cut_mm: float = 0.7):
        self.materials = list(materials)
        self.prod_cut_mm = float(prod_cut_mm)
        names = (C.c_char_p * len(self.materials))(*[m.encode() for m in self.materials])
        d, p = C.c_void_p(), C.c_void_p()
        check("hepem_tables_build", lib().hepem_tables_build(names, len(self.materials), self.prod_cut_mm,
                                                              C.byref(d), C.byref(p)))
        self.data, self.pars = d, p

    @classmethod
    def from_json(cls, text) -> "Tables":
        """Tables from a G4HepEmState document in the reference's JSON schema
        (G4HepEmStateToJson); `text` is str, bytes, or a path to such a file.
        A G4HepEmData document plus a G4HepEmParameters one can be given as a
        pair (data_text, parameters_text)."""
        import os
        if isinstance(text, tuple):
            data_t, pars_t = (t.encode() if isinstance(t, str) else bytes(t) for t in text)
            text = b'{"fData":' + data_t + b',"fParameters":' + pars_t + b"}"
        elif isinstance(text, (str, os.PathLike)) and not str(text).lstrip().startswith("{"):
            with open(text, "rb") as f:
                text = f.read()
        elif isinstance(text, str):
            text = text.encode()
        t = cls.__new__(cls)
        t.materials, t.prod_cut_mm = None, None
        d, p = C.c_void_p(), C.c_void_p()
        check("hepem_tables_from_json", lib().hepem_tables_from_json(text, len(text), C.byref(d), C.byref(p)))
        t.data, t.pars = d, p
        return t

    def set_nuclear(self, enuc_scale: float, gnuc_scale: float) -> "Tables":
        """Fill the e-/e+ nuclear and gamma-nuclear cross sections (scales in
        units of 1/X0 times a fixed energy shape; 0 clears them), so that the
        nuclear branches of the stepping are reached. Returns self."""
        check("hepem_tables_set_nuclear", lib().hepem_tables_set_nuclear(self.data, float(enuc_scale), float(gnuc_scale)))
        return self

    def to_json(self) -> str:
        """The tables as a G4HepEmState document in the reference's JSON schema."""
        size = C.c_uint64()
        check("hepem_tables_to_json", lib().hepem_tables_to_json(self.data, self.pars, None, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        check("hepem_tables_to_json", lib().hepem_tables_to_json(self.data, self.pars, buf, C.byref(size)))
        return buf.raw[:size.value].decode()

    def serialize(self) -> bytes:
        size = C.c_uint64()
        check("hepem_tables_serialize", lib().hepem_tables_serialize(self.data, self.pars, None, C.byref(size)))
        buf = (C.c_uint8 * size.value)()
        check("hepem_tables_serialize", lib().hepem_tables_serialize(self.data, self.pars, buf, C.byref(size)))
        return bytes(buf)

    def digest(self) -> str:
        import hashlib
        return hashlib.sha256(self.serialize()).hexdigest()

    def close(self):
        if self.data:
            lib().hepem_tables_free(self.data, self.pars)
            self.data = self.pars = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
