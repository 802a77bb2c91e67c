"""CPU checks of the C-ABI library: it loads, exports every entry point of
include/arma_hip.h, its host-side helpers (mt.dat I/O, dynamic creation of
Mersenne Twister parameters) are correct, and device entry points fail
cleanly without a GPU.  No compute kernels run here.
"""
import os
import re

import numpy as np
import pytest

import arma_amd
from arma_amd import _abi
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "arma_hip.h")).read()
    declared = set(re.findall(r"\b(arma_hip_[a-z_]+)\s*\(", hdr))
    assert declared == set(_abi.EXPORTS)
    L = arma_amd.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_struct_layouts_match_header():
    import ctypes as C
    assert C.sizeof(_abi.MtParams) == 56
    assert C.sizeof(_abi.MtParams) == C.sizeof(O.MtParams)
    # arma_ar_desc with LP64 alignment padding
    assert C.sizeof(_abi.ArDesc) == 120
    assert C.sizeof(_abi.PotDesc) == 96


def gf2_berlekamp_massey(bits):
    """Connection polynomial of a bit list as a Python int (bit i = x^i)."""
    C, B, L, m = 1, 1, 0, 1
    for n, s in enumerate(bits):
        d = s
        for i in range(1, L + 1):
            d ^= ((C >> i) & 1) & bits[n - i]
        if d == 0:
            m += 1
        elif 2 * L <= n:
            T = C
            C ^= B << m
            L = n + 1 - L
            B = T
            m = 1
        else:
            C ^= B << m
            m += 1
    return C, L


def gf2_mulmod(a, b, f, deg):
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if (a >> deg) & 1:
            a ^= f
    return r


def test_dcmt_parameters_are_primitive():
    """The creator's p=521 records have full period: the output bit stream
    of the oracle generator has linear complexity 521 and its minimal
    polynomial satisfies x^(2^521) = x, which for the prime 2^521-1 means it
    is primitive."""
    for id_, seed in [(0, 1234), (7, 99), (4095, 2024)]:
        p = arma_amd.dcmt_create(521, id_, seed)
        assert (p.aaa & 0xFFFF) == id_ and (p.aaa >> 31) == 1
        assert (p.nn, p.mm, p.rr) == (17, 8, 23)
        assert p.umask == 0xFF800000 and p.lmask == 0x007FFFFF
        draws = O.mt_fill(O.MtParams(*p.as_tuple()), 5489, 2 * 521 + 64)
        bits = [int(v) & 1 for v in draws]
        conn, L = gf2_berlekamp_massey(bits)
        assert L == 521
        assert conn & 1 and (conn >> 521) & 1
        x = 2
        for _ in range(521):
            x = gf2_mulmod(x, x, conn, 521)
        assert x == 2


def test_mt_dat_roundtrip(tmp_path):
    arr = arma_amd.dcmt_create_many(521, 5, 77, 2)
    path = str(tmp_path / "mt.dat")
    arma_amd.write_mt_config(path, arr)
    # 72-byte header + 17 state words per record
    assert os.path.getsize(path) == 5 * (72 + 17 * 4)
    back = arma_amd.read_mt_config(path, 5)
    for a, b in zip(arr, back):
        assert a.as_tuple() == b.as_tuple()
    with pytest.raises(arma_amd.PRNGError):
        arma_amd.read_mt_config(path, 6)


def test_mt_dat_header_layout(tmp_path):
    p = arma_amd.mt19937_params()
    arr = _abi.params_array([p])
    path = str(tmp_path / "one.dat")
    arma_amd.write_mt_config(path, arr)
    raw = np.fromfile(path, dtype=np.uint32)
    assert raw.size == 18 + 624
    assert raw[0] == 0x9908B0DF and raw[1] == 397 and raw[2] == 624 and raw[3] == 31
    assert raw[12] == 0x9D2C5680 and raw[13] == 0xEFC60000


def test_missing_mt_dat():
    with pytest.raises(arma_amd.ArmaHipError):
        arma_amd.read_mt_config("/nonexistent/mt.dat", 1)


def test_device_entry_points_fail_cleanly_without_context():
    import ctypes as C
    L = arma_amd.lib()
    rc = L.arma_hip_synchronize(None)
    assert rc == _abi.ARMA_HIP_EINVAL
    assert b"null context" in L.arma_hip_last_error()


def test_shipped_dcmt_table():
    """The shipped aaa table is what the creator makes for ids 0..4095 with
    the table seed, and its records are primitive."""
    tab = arma_amd._abi.dcmt_p521_params(4096)
    for i in (0, 1, 777, 4095):
        again = arma_amd.dcmt_create(521, i, arma_amd._abi.DCMT_TABLE_SEED)
        assert tab[i].as_tuple() == again.as_tuple()
        assert (tab[i].aaa & 0xFFFF) == i
    p = tab[2024]
    bits = [int(v) & 1 for v in O.mt_fill(O.MtParams(*p.as_tuple()), 1, 2 * 521 + 16)]
    conn, L = gf2_berlekamp_massey(bits)
    assert L == 521
    x = 2
    for _ in range(521):
        x = gf2_mulmod(x, x, conn, 521)
    assert x == 2
