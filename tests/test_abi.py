"""The C-ABI library loads and exports every symbol its headers declare (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from libuw12_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ('uw12g.h', 'uw12g_testing.h'):
        src = open(os.path.join(ROOT, 'include', h)).read()
        src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
        for m in re.finditer(r'\b(uw12g_[a-z0-9_]+)\s*\(', src):
            names.add(m.group(1))
    return names


def test_headers_declare_expected_entry_points():
    names = declared_symbols()
    for n in ('uw12g_setup_base_integrals', 'uw12g_setup_abs_projectors', 'uw12g_energy', 'uw12g_fock',
              'uw12g_form_fock', 'uw12g_last_error'):
        assert n in names
    assert names == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for n in declared_symbols():
        assert hasattr(L, n), n


def test_two_level_map_fast_division():
    L = _lib.load()
    rng = np.random.default_rng(0)
    divisors = [1, 2, 3, 5, 7, 16, 80, 320, 960, 1600, 4500, 65535, 65537, 1 << 20, 3 * (1 << 21) + 7]
    for d in divisors:
        m = _lib.Map2(3, 1000003, d)
        xs = np.concatenate([np.arange(0, 2048), rng.integers(0, 2**31, 4096), [2**32 - 1, 2**32 - 2, 2**31]])
        for x in xs:
            x = int(x)
            assert L.uw12g_test_map2(m, x) == (x % d) * 3 + (x // d) * 1000003, (d, x)


def test_one_level_map():
    L = _lib.load()
    assert L.uw12g_test_map2(_lib.Map2(7, 0, 0), 12345) == 7 * 12345


def test_context_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    L = _lib.load()
    h = _lib.vp()
    rc = L.uw12g_context_create(0, ctypes.byref(h))
    assert rc != 0
    assert L.uw12g_last_error()


def test_time_gemm_refuses_descriptors_past_the_arrays():
    """uw12g_time_gemm checks every operand's extent on the host before it touches a device."""
    L = _lib.load()
    lin = lambda s: _lib.Map2(s, 0, 0)
    a = np.zeros(16)
    p = a.ctypes.data
    ms = ctypes.c_double()
    M, N, K = 8, 6, 5

    def call(nA, nB, nC, Ak=lin(M), batch=1, Ab=0):
        return L.uw12g_time_gemm(0, M, N, K, batch, 1.0, 0.0, p, nA, lin(1), Ak, Ab, p, nB, lin(N), lin(1), 0,
                                 p, nC, lin(1), lin(M), 0, None, 0, lin(0), 0, 1, ctypes.byref(ms))

    for args in ((M * K - 1, K * N, M * N), (M * K, K * N - 1, M * N), (M * K, K * N, M * N - 1),
                 (M * K, K * N, M * N, _lib.Map2(M, 5 * M, 3)), (2 * M * K - 1, K * N, M * N, lin(M), 2, M * K)):
        assert call(*args) == -2, args
        assert b'reaches past' in L.uw12g_last_error()
