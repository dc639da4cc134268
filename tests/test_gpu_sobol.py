"""GPU parity of the scrambled Sobol generator and the inverse CDF (through the C ABI).

The reference generator is kernel_gen_hh_scrambled_sobol_normal
(methods_cuda.cu:753-825).  The engine's uint32 points must be bit-exact with the
oracle's per-point Hong-Hickernell scramble of the Gray-code Sobol point.
"""
import numpy as np
import pytest

import fdmc
import pyoracle as O

pytestmark = pytest.mark.gpu


def _gen_u32(engine, torch, salt, D, n, offset):
    out = torch.empty((D, n), dtype=torch.int32, device="cuda")
    engine.sobol_scrambled_u32(salt, D, n, offset, out.data_ptr())
    return out.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("salt,D,n,offset", [
    (42, 8, 4096, 0),            # first replica of run_qmc_cuda (salt 42 + r)
    (73, 256, 1000, 12345),      # C2 dimension, unaligned offset and count
    (42 + 5 * 16 + 3, 64, 513, 1 << 20),  # MLQMC salt 42 + l*R + r
    (0xFFFFFFFF, 3, 300, 0xFFFFFFFF - 299),  # last representable indices
])
def test_scrambled_u32_bit_exact(engine, torch_cuda, salt, D, n, offset):
    got = _gen_u32(engine, torch_cuda, salt, D, n, offset)
    # oracle: linear-table generator, cross-checked below against the per-point scramble
    ref = O.sobol_scrambled_block(salt, D, offset, n)
    assert np.array_equal(got, ref)
    rng = np.random.default_rng(salt & 0xFFFF)
    for _ in range(64):
        d, p = int(rng.integers(D)), int(rng.integers(n))
        assert got[d, p] == O.sobol_scrambled(salt, d, offset + p)


def test_scrambled_high_dimensions(engine, torch_cuda):
    D, n, off, salt = 20000, 16, 777, 99
    got = _gen_u32(engine, torch_cuda, salt, D, n, off)
    for d in (0, 1, 2, 1023, 1024, 4095, 12345, 19998, 19999):
        for p in (0, 7, 15):
            assert got[d, p] == O.sobol_scrambled(salt, d, off + p), (d, p)


def test_offset_range_rejected(engine):
    with pytest.raises(fdmc.FdmcError) as e:
        engine.sobol_scrambled_u32(1, 4, 10, (1 << 32) - 5, 0xDEAD)  # never launched: rejected on host
    assert e.value.code == 3


def test_normals_vs_fp64_oracle(engine, torch_cuda):
    torch = torch_cuda
    salt, D, n = 42, 32, 8192
    out = torch.empty((D, n), dtype=torch.float32, device="cuda")
    engine.sobol_normal(salt, D, n, 0, out.data_ptr())
    got = out.cpu().numpy()
    u32 = O.sobol_scrambled_block(salt, D, 0, n)
    ref = np.vectorize(O.lib().orc_normal_u32)(u32)
    # FP32 inverse CDF: |err| <= 4e-7 * max(1, |x|) over the whole range
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() < 4e-7, err.max()


def test_icdf_extremes_and_sweep(engine, torch_cuda):
    torch = torch_cuda
    s = np.concatenate([
        np.array([0, 1, 2, 3, 255, 1 << 20, (1 << 31) - 1, 1 << 31, (1 << 31) + 1,
                  (1 << 32) - 256, (1 << 32) - 2, (1 << 32) - 1], np.uint64),
        np.linspace(0, 2 ** 32 - 1, 200001).astype(np.uint64),
        (np.logspace(0, 31.9, 5000, base=2.0)).astype(np.uint64),
        (2 ** 32 - 1 - np.logspace(0, 31.9, 5000, base=2.0)).astype(np.uint64),
    ]).astype(np.uint32)
    d_in = torch.from_numpy(s.view(np.int32).copy()).cuda()
    d_out = torch.empty(s.size, dtype=torch.float32, device="cuda")
    engine.icdf_u32(d_in.data_ptr(), s.size, d_out.data_ptr())
    got = d_out.cpu().numpy().astype(np.float64)
    ref = np.array([O.lib().orc_normal_u32(int(x)) for x in s])
    assert np.all(np.isfinite(got))
    err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
    assert err.max() < 4e-7, (err.max(), s[np.argmax(err)])
    # symmetry of the folded map: x(s) = -x(~s), exact for small integers; at
    # u = 1/2 -+ 2^-33 both values are within 2^-31 of 0
    for i, j in ((0, 11), (1, 10), (4, 9)):
        assert got[i] == -got[j], (i, j)
    assert abs(got[6]) < 1e-9 and abs(got[7]) < 1e-9


def test_scrambled_u32_against_numpy_and_scipy(engine, torch_cuda):
    """The device points equal the numpy restatement of the seeding and scramble
    (methods_cuda.cu:753-795, 819) applied to scipy's independent unscrambled
    Sobol points: no oracle code on this path."""
    from scipy.stats import qmc
    from test_hh_seeding import SHIFT_KEY, _np_scramble, sm

    D, m = 24, 12
    X = (qmc.Sobol(D, scramble=False, bits=32).random_base2(m) * 2.0 ** 32).astype(np.uint64).astype(np.uint32)
    for salt in (42, 42 + 7 * 16 + 5):
        got = _gen_u32(engine, torch_cuda, salt, D, 1 << m, 0)
        for d in range(D):
            base = sm(salt) ^ d
            want = _np_scramble(X[:, d], [sm(base ^ k) for k in range(32)], sm(base ^ SHIFT_KEY))
            assert np.array_equal(got[d], want), (salt, d)
