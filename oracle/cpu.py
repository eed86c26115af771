"""ctypes binding of the C++ CPU oracle (``liboracle_uw12.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the shipped library.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, 'liboracle_uw12.so')
_lib = None

TERMS = ('e2_direct', 'e2_indirect', 'e3_direct', 'e3_indirect', 'e4_direct', 'e4_indirect')


def build():
    subprocess.check_call(['make', '-s', '-C', _HERE, 'liboracle_uw12.so'])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        _lib.oracle_uw12.restype = ctypes.c_int
        _lib.oracle_uw12.argtypes = [dp] * 9 + [ctypes.c_int] * 3 + [dp, ctypes.c_int, ctypes.c_int, dp, ip, ip,
                                                                     ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                                     dp, dp]
        _lib.oracle_last_error.restype = ctypes.c_char_p
        _lib.oracle_setup.restype = ctypes.c_int
        _lib.oracle_setup.argtypes = [dp] * 9 + [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_void_p)]
        _lib.oracle_destroy.argtypes = [ctypes.c_void_p]
        _lib.oracle_destroy.restype = None
        _lib.oracle_evaluate.restype = ctypes.c_int
        _lib.oracle_evaluate.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int, dp, ip, ip, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, dp, dp]
        _lib.oracle_set_pair_stride.argtypes = [ctypes.c_int]
        _lib.oracle_set_pair_stride.restype = None
        _lib.oracle_set_progress.argtypes = [ctypes.c_int]
        _lib.oracle_set_progress.restype = None
        _lib.oracle_abs_projector.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp]
        _lib.oracle_eig_sym.argtypes = [ctypes.c_int, dp, dp, dp]
        _lib.oracle_blas_kind.restype = ctypes.c_int
        _lib.oracle_force_builtin_blas.argtypes = [ctypes.c_int]
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f(a):
    return None if a is None else np.asfortranarray(a, dtype=np.float64)


def uw12(ints, S, C, occ, n_active, k0, k1, indirect=None, fock=True):
    """Evaluate the oracle.  ``ints`` = {'W': (X3, X2, X3ri), 'V': ..., 'WV': (X3, X2, None)}
    (column-major arrays as in the reference), C = list of (n_ao, n_orb) per spin.
    Returns (terms dict, list of Fock matrices or None)."""
    L = lib()
    n_spin = len(C)
    N = C[0].shape[0]
    n_orb = C[0].shape[1]
    X3w, X2w, X3riw = (_f(x) for x in ints['W'])
    X3v, X2v, X3riv = (_f(x) for x in ints['V'])
    X3wv, X2wv, _ = (_f(x) for x in ints['WV'])
    n_df = X2w.shape[0]
    n_ri = X3riw.shape[0] // N
    Cm = np.asfortranarray(np.concatenate([np.asarray(c, dtype=np.float64) for c in C], axis=1))
    occv = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in occ]))
    n_occ = (ctypes.c_int * n_spin)(*[len(x) for x in occ])
    n_act = (ctypes.c_int * n_spin)(*[int(x) for x in n_active])
    if indirect is None:
        indirect = abs(k1) > 4 * np.finfo(float).eps
    Sm = _f(S)
    terms = np.zeros(6)
    F = np.zeros((N, N * n_spin), order='F') if fock else None
    rc = L.oracle_uw12(_p(X3w), _p(X2w), _p(X3riw), _p(X3v), _p(X2v), _p(X3riv), _p(X3wv), _p(X2wv), _p(Sm),
                       N, n_df, n_ri, _p(Cm), n_orb, n_spin, _p(occv), n_occ, n_act, float(k0), float(k1),
                       int(bool(indirect)), _p(F) if fock else None, _p(terms))
    if rc != 0:
        raise RuntimeError(L.oracle_last_error().decode())
    out = dict(zip(TERMS, terms.tolist()))
    out['e2'] = out['e2_direct'] + out['e2_indirect']
    out['e3'] = out['e3_direct'] + out['e3_indirect']
    out['e4'] = out['e4_direct'] + out['e4_indirect']
    out['total'] = out['e2'] + out['e3'] + out['e4']
    Fl = [np.array(F[:, s * N:(s + 1) * N]) for s in range(n_spin)] if fock else None
    return out, Fl


class System:
    """Set-up oracle inputs (DF metric eigendecompositions and S^+ done once)."""

    def __init__(self, ints, S):
        L = lib()
        self._keep = []
        X3w, X2w, X3riw = (_f(x) for x in ints['W'])
        X3v, X2v, X3riv = (_f(x) for x in ints['V'])
        X3wv, X2wv, _ = (_f(x) for x in ints['WV'])
        Sm = _f(S)
        self._keep = [X3w, X2w, X3riw, X3v, X2v, X3riv, X3wv, X2wv, Sm]
        self.n_ao = int(round((np.sqrt(8 * X3w.shape[0] + 1) - 1) / 2))
        n_df = X2w.shape[0]
        n_ri = X3riw.shape[0] // self.n_ao
        h = ctypes.c_void_p()
        rc = L.oracle_setup(_p(X3w), _p(X2w), _p(X3riw), _p(X3v), _p(X2v), _p(X3riv), _p(X3wv), _p(X2wv), _p(Sm),
                            self.n_ao, n_df, n_ri, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(L.oracle_last_error().decode())
        self._h = h

    def evaluate(self, C, occ, n_active, k0, k1, indirect=None, fock=True):
        L = lib()
        n_spin = len(C)
        N = self.n_ao
        n_orb = max(c.shape[1] for c in C)
        Cm = np.zeros((N, n_orb * n_spin), order='F')
        for s, c in enumerate(C):
            Cm[:, s * n_orb:s * n_orb + c.shape[1]] = c
        occv = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in occ]))
        n_occ = (ctypes.c_int * n_spin)(*[len(x) for x in occ])
        n_act = (ctypes.c_int * n_spin)(*[int(x) for x in n_active])
        if indirect is None:
            indirect = abs(k1) > 4 * np.finfo(float).eps
        terms = np.zeros(6)
        F = np.zeros((N, N * n_spin), order='F') if fock else None
        rc = L.oracle_evaluate(self._h, _p(Cm), n_orb, n_spin, _p(occv), n_occ, n_act, float(k0), float(k1),
                               int(bool(indirect)), _p(F) if fock else None, _p(terms))
        if rc != 0:
            raise RuntimeError(L.oracle_last_error().decode())
        out = dict(zip(TERMS, terms.tolist()))
        out['e2'] = out['e2_direct'] + out['e2_indirect']
        out['e3'] = out['e3_direct'] + out['e3_indirect']
        out['e4'] = out['e4_direct'] + out['e4_indirect']
        out['total'] = out['e2'] + out['e3'] + out['e4']
        Fl = [np.array(F[:, s * N:(s + 1) * N]) for s in range(n_spin)] if fock else None
        return out, Fl

    def __del__(self):
        try:
            lib().oracle_destroy(self._h)
        except Exception:
            pass


def set_pair_stride(stride):
    """Evaluate only every stride-th outer same-spin pair index (CPU-baseline sampling)."""
    lib().oracle_set_pair_stride(int(stride))


def set_progress(on):
    """Print completed same-spin pair-loop iterations to stderr (long evaluations)."""
    lib().oracle_set_progress(1 if on else 0)


def abs_projector(S, n_ao, n_ri):
    out = np.zeros((n_ao + n_ri, n_ao + n_ri), order='F')
    rc = lib().oracle_abs_projector(_p(_f(S)), n_ao, n_ri, _p(out))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out


def eig_sym(A):
    A = _f(A)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), order='F')
    rc = lib().oracle_eig_sym(n, _p(A), _p(w), _p(V))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return w, V


def blas_kind():
    return 'mkl' if lib().oracle_blas_kind() == 0 else 'builtin'


def force_builtin_blas(on=True):
    lib().oracle_force_builtin_blas(1 if on else 0)
