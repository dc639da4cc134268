"""Independent pins of the Hong-Hickernell seeding (no GPU).

The reference scrambles Sobol coordinate `dim` of replica `salt` with
(methods_cuda.cu:753-759, 781-795, 819):
    base  = splitmix32(salt) ^ dim
    row k = (1 << (31 - k)) | (splitmix32(base ^ k) & bits above 31 - k)
    bit (31 - k) of the output = parity(row k & raw)
    shift = splitmix32(base ^ 0xA5A5A5A5), XORed into the output
splitmix32 (methods_cuda.cu:753-759) adds the golden-ratio increment 0x9E3779B9
and applies the two-round xor-shift-multiply hash [16 21f0aaad 15 735a2d97 15]
of Wellons' hash-prospector tables.

The oracle's generator is checked here by reading its matrix back from its own
points (columns L e_b from the scrambled direction words and the unscrambled
direction numbers), never by retyping its code path:
  * splitmix32 against two separate implementations (numpy uint32 arrays and
    Python integers) of the published hash;
  * the digital shift equals the scrambled image of the raw point 0, and that
    image equals splitmix32(splitmix32(salt) ^ dim ^ 0xA5A5A5A5);
  * column b of the matrix has bit b set, nothing above it, and bit c < b equal
    to bit b of splitmix32(splitmix32(salt) ^ dim ^ (31 - c));
  * row seeds alias across dimensions exactly where dim ^ k = dim' ^ k' (a
    property of base ^ k that the oracle's rows must show with no formula);
  * the whole point chain against a numpy restatement fed by scipy's
    independent unscrambled Sobol generator.
Each check is also run on planted mistakes (salt ^ dim inside the hash, base + k
for base ^ k, a wrong shift constant, a wrong hash constant, rows indexed from
the other end) to show that it would catch them.
"""
import numpy as np
import pytest
from scipy.stats import qmc

import pyoracle as O

U32 = np.uint32
SHIFT_KEY = 0xA5A5A5A5


# ------------------------------------------------------------ independent hashes
def sm32_np(x, c1=0x21F0AAAD, c2=0x735A2D97):
    """splitmix32 on numpy uint32 arrays (multiplication wraps mod 2^32)."""
    x = np.asarray(x, dtype=U32) + U32(0x9E3779B9)
    x = (x ^ (x >> U32(16))) * U32(c1)
    x = (x ^ (x >> U32(15))) * U32(c2)
    return x ^ (x >> U32(15))


def sm32_int(x):
    """splitmix32 with Python integers reduced mod 2^32 after every operation."""
    m = 0xFFFFFFFF
    x = (x + 0x9E3779B9) & m
    x = ((x ^ (x >> 16)) * 0x21F0AAAD) & m
    x = ((x ^ (x >> 15)) * 0x735A2D97) & m
    return x ^ (x >> 15)


def sm(x):
    return int(sm32_np(np.array([x & 0xFFFFFFFF], dtype=U32))[0])


# ------------------------------------------------------------ generators under test
def oracle_point(salt, dim, index):
    return int(O.sobol_scrambled(salt, dim, index))


def _np_scramble(raw, row_seeds, shift):
    out = np.zeros_like(raw)
    for k in range(32):
        mask = 0 if k == 0 else (0xFFFFFFFF << (32 - k)) & 0xFFFFFFFF
        row = U32(1 << (31 - k)) | (U32(row_seeds[k]) & U32(mask))
        par = np.zeros_like(raw)
        v = raw & row
        for b in range(32):
            par ^= (v >> U32(b)) & U32(1)
        out |= par << U32(31 - k)
    return out ^ U32(shift)


def make_numpy_point(base_fn=None, row_fn=None, shift_key=SHIFT_KEY, hash_fn=None):
    """Numpy generator of the scrambled point with pluggable pieces (for the
    planted mistakes); the defaults follow the reference's formulas."""
    h = hash_fn or sm
    base_fn = base_fn or (lambda salt, dim: h(salt) ^ dim)
    row_fn = row_fn or (lambda base, k: h(base ^ k))

    def point(salt, dim, index):
        v = O.dirvectors(dim)
        g = index ^ (index >> 1)
        raw = 0
        for j in range(32):
            if (g >> j) & 1:
                raw ^= int(v[j])
        base = base_fn(salt, dim) & 0xFFFFFFFF
        seeds = [row_fn(base, k) & 0xFFFFFFFF for k in range(32)]
        return int(_np_scramble(np.array([raw], dtype=U32), seeds, h(base ^ shift_key))[0])

    return point


MUTANTS = {
    "hash of salt^dim": dict(base_fn=lambda s, d: sm(s ^ d)),
    "base+k for base^k": dict(row_fn=lambda b, k: sm((b + k) & 0xFFFFFFFF)),
    "shift constant 0x5A5A5A5A": dict(shift_key=0x5A5A5A5A),
    "rows indexed LSB first": dict(row_fn=lambda b, k: sm(b ^ (31 - k))),
    "second hash constant": dict(hash_fn=lambda x: int(sm32_np(np.array([x & 0xFFFFFFFF], dtype=U32),
                                                                c2=0x735A2D95)[0])),
}


# ------------------------------------------------------------ matrix read-out
def columns(point, salt, dim):
    """Columns L e_b (b = bit position) of the scramble matrix of (salt, dim),
    read from the generator's points: the scrambled direction word of v_j is
    point((2 << j) - 1) ^ point(0) (Gray code of (2 << j) - 1 is 1 << j); v_j has
    its lowest set bit at 31 - j, so every unit vector is a XOR of v_j's found by
    clearing bits from the bottom up."""
    v = [int(x) for x in O.dirvectors(dim)]
    e = point(salt, dim, 0)
    Lv = [point(salt, dim, (2 << j) - 1) ^ e if j < 31 else point(salt, dim, 0xFFFFFFFF) ^ e for j in range(32)]
    cols = []
    for b in range(32):
        x, img = 0, 0
        want = 1 << b
        # combine v's until x == want: clear the lowest differing bit each time
        while x != want:
            diff = x ^ want
            c = (diff & -diff).bit_length() - 1
            j = 31 - c
            x ^= v[j]
            img ^= Lv[j]
        cols.append(img)
    return cols, e


def check_shift(point, salt, dim):
    return point(salt, dim, 0) == sm(sm(salt) ^ dim ^ SHIFT_KEY)


def check_columns(point, salt, dim):
    cols, _ = columns(point, salt, dim)
    base = sm(salt) ^ dim
    for b, col in enumerate(cols):
        if (col >> b) != 1:          # diagonal set, nothing above it
            return False
        for c in range(b):
            if ((col >> c) & 1) != ((sm(base ^ (31 - c)) >> b) & 1):
                return False
    return True


def rows_from_columns(cols):
    """Row k (output bit 31 - k) as the word of its entries over the input bits."""
    return [sum(((cols[b] >> (31 - k)) & 1) << b for b in range(32)) for k in range(32)]


def check_aliasing(point, salt, d1, d2):
    """Rows k1 of d1 and k2 of d2 with d1 ^ k1 = d2 ^ k2 share their row seed, so
    they agree on every input bit above both diagonals; rows of pairs that do not
    alias agree on those bits only by chance (checked to differ somewhere)."""
    r1 = rows_from_columns(columns(point, salt, d1)[0])
    r2 = rows_from_columns(columns(point, salt, d2)[0])
    shared = 0
    differ = 0
    for k1 in range(32):
        for k2 in range(32):
            lo = max(31 - k1, 31 - k2) + 1
            if lo > 31:
                continue
            mask = ((1 << 32) - 1) ^ ((1 << lo) - 1)
            same = (r1[k1] & mask) == (r2[k2] & mask)
            if d1 ^ k1 == d2 ^ k2:
                if not same:
                    return False
                shared += 1
            elif lo <= 24 and not same:
                differ += 1
    return shared > 0 and differ > 0


# ------------------------------------------------------------ tests
def test_splitmix32_known_answers():
    rng = np.random.default_rng(20260)
    xs = np.concatenate([rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(U32),
                         np.array([0, 1, 42, 0x9E3779B9, 0x61C88647, 0xA5A5A5A5, 0xFFFFFFFF], dtype=U32)])
    ref = sm32_np(xs)
    for x, r in zip(xs[-7:].tolist(), ref[-7:].tolist()):
        assert sm32_int(x) == r
    got = np.array([O.lib().orc_splitmix32(int(x)) for x in xs], dtype=U32)
    assert np.array_equal(got, ref)
    # one worked value: x = 0 -> 0x9E3779B9 -> the two xor-shift-multiply rounds
    assert sm32_int(0) == int(ref[-7])


@pytest.mark.parametrize("salt,dim", [(42, 0), (42, 1), (43, 255), (42 + 5 * 16 + 3, 17), (2026, 511),
                                      (0xFFFFFFFF, 19999), (0, 1000)])
def test_shift_is_image_of_zero(salt, dim):
    Lv, e = O.scrambled_dirs(salt, dim)
    assert e == oracle_point(salt, dim, 0)
    assert check_shift(oracle_point, salt, dim)


@pytest.mark.parametrize("salt,dim", [(42, 0), (42, 3), (73, 256), (42 + 31, 1023), (123456789, 19999)])
def test_matrix_columns_follow_row_seeds(salt, dim):
    assert check_columns(oracle_point, salt, dim)


@pytest.mark.parametrize("salt,d1,d2", [(42, 0, 1), (42, 2, 7), (99, 100, 117), (2026, 5, 30)])
def test_row_seeds_alias_across_dimensions(salt, d1, d2):
    assert check_aliasing(oracle_point, salt, d1, d2)


def test_point_chain_against_numpy_and_scipy():
    """Scrambled points = numpy scramble of scipy's unscrambled Sobol points."""
    rng = np.random.default_rng(7)
    n = 1 << 12
    D = 40
    X = (qmc.Sobol(D, scramble=False, bits=32).random_base2(12) * 2.0 ** 32).astype(np.uint64).astype(U32)
    for salt in (42, 42 + 31, 0xFFFFFFFF):
        for d in rng.choice(D, 6, replace=False).tolist():
            base = sm(salt) ^ d
            seeds = [sm(base ^ k) for k in range(32)]
            want = _np_scramble(X[:, d], seeds, sm(base ^ SHIFT_KEY))
            got = O.sobol_scrambled_block(salt, D, 0, n)[d]
            assert np.array_equal(got, want), (salt, d)


def test_numpy_restatement_passes_every_check():
    p = make_numpy_point()
    assert check_shift(p, 42, 3) and check_columns(p, 42, 3) and check_aliasing(p, 42, 2, 7)
    for idx in (0, 1, 12345, 0xFFFFFFFF):
        assert p(42, 3, idx) == oracle_point(42, 3, idx)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistakes_are_caught(name):
    p = make_numpy_point(**MUTANTS[name])
    caught = [not check_shift(p, 42, 3), not check_columns(p, 42, 3), not check_aliasing(p, 42, 2, 7)]
    assert any(caught), name


def test_golden_vectors_match_numpy_restatement():
    """The committed regression vectors agree with the numpy restatement (they
    were written from the oracle; this makes them independently checked)."""
    import json
    import os

    p = make_numpy_point()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sobol_scrambled.json")
    for v in json.load(open(path)):
        assert p(v["salt"], v["dim"], v["index"]) == v["value"], v
