"""RANECU and its skip-ahead (MCGPU-PET_kernel.cu:898-1083, MCGPU-PET.cu:2906-2935), pinned by
integer arithmetic done independently in Python, and the library's host-side jump-table
skip-ahead checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

import mcpet
import oracle_py

M1, A1, M2, A2 = 2147483563, 40014, 2147483399, 40692


def test_abmodm_brute_force():
    L = oracle_py.load("libm")
    rng = np.random.default_rng(1)
    for m in (M1, M2):
        a = rng.integers(1, m, 20000)
        s = rng.integers(1, m, 20000)
        edge = [1, 2, 32767, 32768, 32769, 65535, m - 1, m - 2, 40014, 40692, 1 << 30]
        for x in edge:
            for y in edge:
                assert L.orc_abMODm(m, x, y) == (x * y) % m
        for x, y in zip(a.tolist(), s.tolist()):
            assert L.orc_abMODm(m, x, y) == (x * y) % m


def _py_stream(sx, sy, n):
    """RANECU by direct modular recurrence (no Schrage), returning the integer outputs."""
    out = []
    for _ in range(n):
        sx = (A1 * sx) % M1
        sy = (A2 * sy) % M2
        z = sx - sy
        if z < 1:
            z += M1 - 1
        out.append(z)
    return out, sx, sy


def test_ranecu_float_is_exact_power_of_two_scaling():
    # 4.65661305739e-10f rounds to exactly 2^-31 in binary32
    assert np.float32(4.65661305739e-10) == np.float32(2.0 ** -31)
    L = oracle_py.load("libm")
    sx, sy = C.c_int(12345), C.c_int(12345)
    ints, _, _ = _py_stream(12345, 12345, 5000)
    got = np.array([L.orc_ranecu(C.byref(sx), C.byref(sy)) for _ in range(5000)], np.float32)
    want = (np.array(ints, np.float64).astype(np.float32) * np.float32(2.0 ** -31)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got.min() >= np.float32(2.0 ** -31) and got.max() <= np.float32(1.0 - 2.0 ** -24)
    sx, sy = C.c_int(777), C.c_int(777)
    ints, _, _ = _py_stream(777, 777, 2000)
    gd = np.array([L.orc_ranecu_double(C.byref(sx), C.byref(sy)) for _ in range(2000)])
    assert np.array_equal(gd, np.array(ints, np.float64) * 4.6566130573917692e-10)


def test_skip_ahead_equals_sequential_steps():
    """init_PRNG(id, hpt, seed) must land where (id+1)*hpt*256 sequential steps land."""
    L = oracle_py.load("libm")
    for hpt, ids in ((1, [0, 1, 5, 37, 300]), (3, [0, 2, 11])):
        for i in ids:
            sx, sy = C.c_int(), C.c_int()
            L.orc_init_PRNG(i, hpt, 987654, C.byref(sx), C.byref(sy))
            leap = (i + 1) * hpt * 256
            assert sx.value == (987654 * pow(A1, leap, M1)) % M1
            assert sy.value == (987654 * pow(A2, leap, M2)) % M2
            # brute force: leap single steps of each MLCG
            x, y = 987654, 987654
            for _ in range(leap):
                x = (A1 * x) % M1
                y = (A2 * y) % M2
            assert (sx.value, sy.value) == (x, y)


def test_update_seed_composition():
    L = oracle_py.load("libm")
    s = C.c_int(123456789)
    L.orc_update_seed_PRNG(3, 1000, C.byref(s))
    assert s.value == (123456789 * pow(A1, 3 * 1000 * 256, M1)) % M1
    s0 = C.c_int(55555)
    L.orc_update_seed_PRNG(0, 1000, C.byref(s0))
    assert s0.value == 55555


def test_library_jump_table_matches_reference_skip_ahead():
    """The library's host-side seeds (jump table over a^(L 2^k)) equal the reference's
    square-and-multiply init_PRNG for ids across the int32 range and beyond."""
    L = oracle_py.load("libm")
    ids = np.array([0, 1, 2, 3, 63, 64, 1000, 65535, 123456789, 2**31 - 1, 2**31, 3 * 2**33 + 1, 10**11], np.int64)
    rs = np.random.default_rng(9).integers(1, 2**31 - 1, 6).tolist()
    for seed in [1, 12345, 32768, 32769, -3, -M2, M2 - 1, M2, M2 + 1, M1 - 1, M1, 2**31 - 1] + rs:
        sx, sy = mcpet.stream_seeds(seed, ids)
        for k, i in enumerate(ids.tolist()):
            ox, oy = C.c_int(), C.c_int()
            L.orc_init_PRNG(i, 555555, seed, C.byref(ox), C.byref(oy))
            assert (int(sx[k]), int(sy[k])) == (ox.value, oy.value), (seed, i)
            if 0 < seed < M2:
                leap = (i + 1) * 555555 * 256
                assert int(sx[k]) == (seed * pow(A1, leap, M1)) % M1
                assert int(sy[k]) == (seed * pow(A2, leap, M2)) % M2


def test_pseudo_mersenne_mulmod_brute_force():
    """The library's stream set-up reduces products modulo 2^31 - c by folding; check the
    folded product against Python integers over random and extreme operands, through
    the public seed routine: seeds of stream 0 are seed * a^L mod m."""
    rng = np.random.default_rng(21)
    seeds = np.concatenate([rng.integers(1, M2, 3000), [1, 2, M2 - 1, 2**30, 2**31 - 250]]).astype(np.int64)
    L1, L2 = pow(A1, 555555 * 256, M1), pow(A2, 555555 * 256, M2)
    for seed in seeds.tolist():
        sx, sy = mcpet.stream_seeds(int(seed), np.array([0], np.int64))
        assert int(sx[0]) == seed * L1 % M1 and int(sy[0]) == seed * L2 % M2


def test_seeds_the_reference_cannot_use_are_rejected():
    """0 and seeds below -m2 make abMODm divide by zero (K:1000-1032); the library refuses them."""
    for seed in (0, -M2 - 1, -2**31):
        with pytest.raises(mcpet.McpetError):
            mcpet.stream_seeds(seed, np.array([0], np.int64))
