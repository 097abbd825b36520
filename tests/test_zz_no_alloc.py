"""hepem_step_noalloc -- the step with the library's fixed per-GPU work array --
against hepem_step with a caller work buffer, through the C-ABI.

The two run the same kernels, so every output and every field of the updated
track state must agree bit for bit. Two calls in a row, the second on a
smaller batch, check that nothing the first call left in the shared array
leaks into the second. The capacity check is a host-side rejection.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import same_bits, valid_secondaries

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


def _run(ctx, seed, tr, dev, noalloc):
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA
    n = len(tr["ekin"])
    tin = TrackSoA.from_numpy(tr, "torch", dev)
    tout = TrackSoA(n, "torch", dev)
    res = StepOutSoA(n, "torch", dev)
    if noalloc:
        ctx.step_noalloc(seed, tin, tout, res)
    else:
        ctx.step(seed, tin, tout, res, ctx.work_buffer(n, dev))
    torch.cuda.synchronize()
    return tout.to_numpy(), valid_secondaries(res.to_numpy())


@pytest.mark.parametrize("precision", ["fp64", "fp32t"])
def test_noalloc_step_equals_step_twice(tables2, dev, precision):
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, Context
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(tables2, HEPEM_FP64 if precision == "fp64" else HEPEM_FP32T)
    # the full capacity first, then a smaller batch on the same array
    for call, (n, seed, tseed) in enumerate(((262144, 0x5EED2026, 11), (37000, 77, 12))):
        _print(f"noalloc {precision} call {call}: n={n} seed={seed}")
        tr = step_tracks(n, tseed, 4)
        t_ref, r_ref = _run(ctx, seed, tr, dev, noalloc=False)
        t_na, r_na = _run(ctx, seed, tr, dev, noalloc=True)
        for k in r_ref:
            assert same_bits(r_na[k], r_ref[k]), f"{precision} call {call}: result {k}"
        for k in t_ref:
            assert same_bits(t_na[k], t_ref[k]), f"{precision} call {call}: state {k}"
        assert len(np.unique(r_na["winner"])) >= 4
    ctx.close()


def test_noalloc_rejects_batches_beyond_capacity(tables2, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd._native import NOALLOC_MAX_TRACKS, lib
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA
    ctx = Context(tables2, HEPEM_FP64)
    n = NOALLOC_MAX_TRACKS + 1
    tin, tout, res = TrackSoA(n, "torch", dev), TrackSoA(n, "torch", dev), StepOutSoA(n, "torch", dev)
    si, so, rr = tin.struct(), tout.struct(), res.struct()
    s = int(torch.cuda.current_stream().cuda_stream)
    assert lib().hepem_step_noalloc(ctx.h, C.c_uint64(1), C.byref(si), C.byref(so), C.byref(rr), s) == -3
    # an empty batch is a no-op
    e_in, e_out, e_res = TrackSoA(0, "torch", dev), TrackSoA(0, "torch", dev), StepOutSoA(0, "torch", dev)
    ei, eo, er = e_in.struct(), e_out.struct(), e_res.struct()
    assert lib().hepem_step_noalloc(ctx.h, C.c_uint64(1), C.byref(ei), C.byref(eo), C.byref(er), s) == 0
    ctx.close()
