"""The C-ABI library on the CPU: it loads, exports every entry point the
headers declare, and its argument checks return errors instead of faulting.
No kernel is launched here."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hepem_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from g4hepem_amd import lib
    L = lib()
    names = declared_functions()
    assert len(names) >= 27, names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_null_arguments_are_rejected():
    from g4hepem_amd import lib
    from g4hepem_amd._native import StepOut, Tracks
    L = lib()
    h = C.c_void_p()
    assert L.hepem_ctx_create(None, None, 0, C.byref(h)) == -1
    assert L.hepem_ctx_create(None, None, 0, None) == -1
    assert L.hepem_ctx_destroy(None) == -1
    assert L.hepem_lookup_eloss_xsec(None, 0, 10, None, None, None, None, None) == -1
    t, r = Tracks(), StepOut()
    assert L.hepem_step(None, 1, C.byref(t), C.byref(t), C.byref(r), None, None) == -1
    assert L.hepem_step_prepare(None, C.byref(t), None, None) == -1
    assert L.hepem_step_run(None, 1, C.byref(t), C.byref(t), C.byref(r), None, None) == -1
    assert L.hepem_step_noalloc(None, 1, C.byref(t), C.byref(t), C.byref(r), None) == -1
    assert L.hepem_lookup_gamma_xsec(None, 10, None, None, None, None, None, None, None, None) == -1
    assert L.hepem_tracks_reset_state(None, None) == -1
    assert L.hepem_tracks_reset_state(C.byref(t), None) == 0   # empty batch: nothing to do
    t5 = Tracks()
    t5.n = 5
    assert L.hepem_tracks_reset_state(C.byref(t5), None) == -1  # null arrays
    assert L.hepem_step_batched(None, 1, None, 0, None) == -1
    assert L.hepem_step_rejected(None, None, None) == -1
    from g4hepem_amd._native import HowFarState
    hs = HowFarState()
    for name in ("hepem_howfar", "hepem_electron_howfar", "hepem_gamma_howfar"):
        assert getattr(L, name)(None, 1, C.byref(t), C.byref(hs), None, None) == -1, name
    for name in ("hepem_perform", "hepem_electron_perform", "hepem_gamma_perform"):
        assert getattr(L, name)(None, 1, C.byref(t), C.byref(hs), C.byref(r), None, None) == -1, name
    assert L.hepem_batch_plan_create(None, None, 1, None) == -1
    assert L.hepem_batch_plan_run(None, 1, None) == -1
    assert L.hepem_batch_plan_tune(None, 1, None) == -1
    assert L.hepem_batch_plan_destroy(None) == -1
    from g4hepem_amd._native import BatchShape
    bs = BatchShape(256, 4096, 0, 0)
    assert L.hepem_batch_plan_get_shape(None, C.byref(bs)) == -1
    assert L.hepem_batch_plan_set_shape(None, C.byref(bs)) == -1
    assert L.hepem_stack_create(None, 10, C.byref(h)) == -1
    assert L.hepem_stack_create(None, 10, None) == -1
    assert L.hepem_stack_destroy(None) == -1
    assert L.hepem_stack_tracks(None, C.byref(t)) == -1
    assert L.hepem_stack_push(None, C.byref(t), None) == -1
    assert L.hepem_stack_copy_out(None, C.byref(t), None) == -1
    m = C.c_int32()
    assert L.hepem_stack_advance(None, C.byref(t), C.byref(r), None, None, None, None, C.byref(m), None) == -1
    assert L.hepem_ctx_set_lookup_staging(None, 1) == -1
    assert L.hepem_ctx_set_process_sort(None, 1) == -1
    assert L.hepem_status_string(-3) == b"capacity exceeded"
    assert L.hepem_step_work_bytes(1000) >= 4000


def test_table_builder_errors():
    from g4hepem_amd import HepEmError, Tables
    with pytest.raises(HepEmError):
        Tables(["G4_Unobtainium"], 0.7)
    with pytest.raises(HepEmError):
        Tables(["G4_Pb"], -1.0)


def test_ctx_create_without_gpu_or_with(tables2):
    """Without a GPU the context reports no device; with one it is created."""
    from g4hepem_amd import lib
    h = C.c_void_p()
    rc = lib().hepem_ctx_create(tables2.data, tables2.pars, 0, C.byref(h))
    assert rc in (0, -5), rc
    if rc == 0:
        assert lib().hepem_ctx_destroy(h) == 0
    assert lib().hepem_ctx_create(tables2.data, tables2.pars, 7, C.byref(h)) == -1


def test_oracle_library_is_separate():
    """The product library does not link the oracle."""
    from g4hepem_amd._native import LIB_PATH
    data = LIB_PATH.read_bytes()
    assert b"oracle_step" not in data and b"ref_step" not in data


def test_type_ordered_batch():
    """type_ordered keeps every track whole and puts the e-/e+ before the
    gammas, each kind in its original order (the layout of a stack advance)."""
    import numpy as np
    from g4hepem_amd.tracks import TRACK_FIELDS, step_tracks, type_ordered
    n = 1003
    t = step_tracks(n, 5, 4)
    o = type_ordered(t)
    ng = int((t["ptype"] == 2).sum())
    assert (o["ptype"][: n - ng] != 2).all() and (o["ptype"][n - ng:] == 2).all()
    ids = o["id"]
    assert (np.diff(ids[: n - ng]) > 0).all() and (np.diff(ids[n - ng:]) > 0).all()
    pos = {int(v): i for i, v in enumerate(t["id"])}
    src = np.array([pos[int(v)] for v in ids])
    for k, (_, c) in TRACK_FIELDS.items():
        assert np.array_equal(o[k].reshape(c, n), t[k].reshape(c, n)[:, src]), k
