"""fdmc_qmc_accumulate_noalloc through the C ABI: the pass that allocates
nothing (direction records and partials in the library's static device array,
sums into the caller's device buffer) returns exactly the sums of
fdmc_qmc_accumulate on the same inputs, on two calls in a row (the second one
smaller, so stale scratch contents from the first would show), and rejects a
pass larger than its scratch with FDMC_E_RANGE before any device work."""
import ctypes as C

import numpy as np
import pytest

import fdmc

pytestmark = pytest.mark.gpu

GBM3 = fdmc.GBMParams(100.0, 0.05, 0.3, 1.0)
ASIAN, GEO = fdmc.Asian(100.0), fdmc.GeomAsian(100.0)
DUP = fdmc.DupireLocalParams(100.0, 0.05, 0.2, 0.5, 0.7, 1.0)
EURO = fdmc.European(100.0, 0.05, 1.0)


def _noalloc(engine, torch, model, payoff, n_steps, R, salt, offset, n, mode, bb=None, ctrl=None, E=0.0, beta=0.0):
    nacc = 7 if ctrl else 2
    out = torch.full((R * nacc,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    m, p = model.to_c(), payoff.to_c()
    cm = C.byref(ctrl[0].to_c()) if ctrl else None
    cp = C.byref(ctrl[1].to_c()) if ctrl else None
    got_nacc = C.c_int32()
    rc = engine.lib.fdmc_qmc_accumulate_noalloc(engine.ctx, C.byref(m), C.byref(p), cm, cp, E, beta, n_steps, R,
                                                salt, offset, n, mode, bb.handle if bb else None,
                                                C.c_void_p(out.data_ptr()), C.byref(got_nacc))
    if rc != 0:
        raise fdmc.FdmcError(rc, engine.lib.fdmc_last_error().decode())
    assert got_nacc.value == nacc
    return out.cpu().numpy().reshape(R, nacc)


def test_noalloc_pass_equals_accumulate_twice(engine, torch_cuda):
    torch = torch_cuda
    n_steps = 256
    E = fdmc.geom_asian_analytic(100.0, 100.0, 1.0, 0.05, 0.3, n_steps)
    bb = engine.bb_upload(fdmc.bb_precompute(n_steps, 1.0))
    try:
        # first call: C2's pass shape (bridge, geometric control), 16 replicas x 20000 points
        print("call 1: R=16 n=20000 D=256 bridge + control", flush=True)
        got = _noalloc(engine, torch, GBM3, ASIAN, n_steps, 16, 42, 777, 20000, fdmc.NOISE_BB, bb=bb,
                       ctrl=(GBM3, GEO), E=E, beta=1.0)
        ref = np.array(engine.qmc_accumulate(GBM3, ASIAN, n_steps, 16, 42, 777, 20000, fdmc.NOISE_BB, dev_bb=bb,
                                             ctrl_model=GBM3, ctrl_payoff=GEO, E_ctrl=E, beta=1.0))
        assert np.array_equal(got, ref[:, :7]), (got, ref)
    finally:
        bb.free()
    # second call, smaller: raw noise, 3 replicas x 1000 points, no control
    print("call 2: R=3 n=1000 D=64 raw", flush=True)
    got2 = _noalloc(engine, torch, DUP, EURO, 64, 3, 7, 0, 1000, fdmc.NOISE_RAW)
    ref2 = np.array(engine.qmc_accumulate(DUP, EURO, 64, 3, 7, 0, 1000, fdmc.NOISE_RAW))
    assert np.array_equal(got2, ref2[:, :2]), (got2, ref2)
    assert np.all(np.isfinite(got2)) and np.all(got2[:, 0] > 0)


def test_noalloc_rejects_passes_beyond_scratch(engine, torch_cuda):
    torch = torch_cuda
    # 256 replicas x 2048 dates of direction records = 128 MiB > 32 MiB
    with pytest.raises(fdmc.FdmcError) as e:
        _noalloc(engine, torch, GBM3, EURO, 2048, 256, 42, 0, 256, fdmc.NOISE_RAW)
    assert e.value.code == 3
    with pytest.raises(fdmc.FdmcError) as e:     # PCA noise needs GEMM buffers: not offered
        _noalloc(engine, torch, GBM3, EURO, 64, 1, 42, 0, 256, fdmc.NOISE_PCA)
    assert e.value.code == 2
    # the context still works afterwards
    got = _noalloc(engine, torch, GBM3, EURO, 64, 2, 42, 0, 512, fdmc.NOISE_RAW)
    ref = np.array(engine.qmc_accumulate(GBM3, EURO, 64, 2, 42, 0, 512, fdmc.NOISE_RAW))
    assert np.array_equal(got, ref[:, :2])
