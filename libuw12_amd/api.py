"""Python mirror of the libuw12 public API (zwill22/libuw12, libuw12.hpp:12-171).

Same function names, argument meaning, layouts and error cases as the
reference; every call goes through the C ABI of libuw12g.so (include/uw12g.h)
into HIP on one GPU.  Errors the reference raises as std::runtime_error /
std::logic_error raise ``UW12Error`` (a RuntimeError) here.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import UW12Error, check, dp, szp, vp

TERMS = ('e2_direct', 'e2_indirect', 'e3_direct', 'e3_indirect', 'e4_direct', 'e4_indirect')
# uw12g_precision values (include/uw12g.h)
FP64, FP32, MIXED, MIXED_BF16X3, MIXED_BF16X6, MIXED_FP16X3 = 0, 1, 2, 3, 4, 5


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(dp)


class Context:
    """One GPU, one HIP stream, and the device caches (no reference counterpart)."""

    _default = None

    def __init__(self, device=0):
        L = _lib.load()
        h = vp()
        check(L.uw12g_context_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = device

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def set_profiling(self, on=True):
        check(_lib.load().uw12g_context_set_profiling(self._h, 1 if on else 0))

    def set_verbosity(self, level=1):
        """Progress lines on stderr at the major stages of each evaluation (level > 0)."""
        check(_lib.load().uw12g_context_set_verbosity(self._h, int(level)))

    def profile(self, cap=4096):
        """Per-launch records of the last evaluation: list of (name, ms, flops, bytes)."""
        L = _lib.load()
        names = (ctypes.c_char_p * cap)()
        ms = np.zeros(cap)
        fl = np.zeros(cap)
        by = np.zeros(cap)
        n = L.uw12g_context_profile(self._h, names, _ptr(ms), _ptr(fl), _ptr(by), cap)
        n = min(n, cap)
        return [(names[i].decode(), ms[i], fl[i], by[i]) for i in range(n)]

    def close(self):
        if getattr(self, '_h', None):
            _lib.load().uw12g_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BaseIntegrals:
    """Device-resident DF integrals of one operator (base_integrals.hpp:46-293)."""

    def __init__(self, handle, ctx, n_ao, n_df, n_ri):
        self._h = handle
        self.ctx = ctx
        self.n_ao, self.n_df, self.n_ri = n_ao, n_df, n_ri

    @property
    def n_kept(self):
        """DF metric eigenvectors kept by the 1e-10 pseudo-inverse."""
        return int(_lib.load().uw12g_base_integrals_kept(self._h))

    def get_number_ao(self):
        return self.n_ao

    def get_number_df(self):
        return self.n_df

    def get_number_ri(self):
        return self.n_ri

    def __del__(self):
        try:
            if self._h:
                _lib.load().uw12g_base_integrals_destroy(self._h)
                self._h = None
        except Exception:
            pass


class ABSProjectors:
    """S^+ over the AO + RI space (ri_utils.hpp:20-74)."""

    def __init__(self, handle, ctx, n_ao, n_ri):
        self._h = handle
        self.ctx = ctx
        self.n_ao, self.n_ri = n_ao, n_ri

    def matrix(self):
        M = self.n_ao + self.n_ri
        out = np.zeros((M, M), order='F')
        check(_lib.load().uw12g_abs_projectors_get(self._h, _ptr(out)))
        return out

    def __del__(self):
        try:
            if self._h:
                _lib.load().uw12g_abs_projectors_destroy(self._h)
                self._h = None
        except Exception:
            pass


def setup_base_integrals(X3, X2, X3_ri=None, n_ao=None, n_df=None, n_ri=None, copy_data=False, context=None):
    """libuw12.hpp:42-50 (with X3_ri) and :64-70 (without).  Arrays are column-major
    (n_ao(n_ao+1)/2, n_df), (n_df, n_df), (n_ao n_ri, n_df); they are always copied."""
    ctx = context or Context.default()
    X3 = _f64(X3)
    X2 = _f64(X2)
    if n_df is None:
        n_df = X2.shape[0]
    if n_ao is None:
        n_ao = int((np.sqrt(8 * X3.shape[0] + 1) - 1) // 2)
    if X3.shape != (n_ao * (n_ao + 1) // 2, n_df) or X2.shape != (n_df, n_df):
        raise UW12Error(-1, 'Number of density-fitting basis functions inconsistent')
    if X3_ri is not None:
        X3_ri = _f64(X3_ri)
        if n_ri is None:
            n_ri = X3_ri.shape[0] // n_ao
        if X3_ri.shape != (n_ao * n_ri, n_df):
            raise UW12Error(-1, 'X3_ri has inconsistent shape')
    else:
        n_ri = 0
    h = vp()
    check(_lib.load().uw12g_setup_base_integrals(ctx._h, _ptr(X3), _ptr(X2), _ptr(X3_ri), n_ao, n_df, n_ri,
                                                 ctypes.byref(h)))
    return BaseIntegrals(h, ctx, n_ao, n_df, n_ri)


def setup_base_integrals_device(X3_dev, X2, X3_ri_dev, n_ao, n_df, n_ri, context=None):
    """Same as setup_base_integrals for device-resident X3 / X3_ri (integer device addresses)."""
    ctx = context or Context.default()
    X2 = _f64(X2)
    h = vp()
    check(_lib.load().uw12g_setup_base_integrals_device(ctx._h, vp(X3_dev), _ptr(X2),
                                                        vp(X3_ri_dev) if X3_ri_dev else None, n_ao, n_df, n_ri,
                                                        ctypes.byref(h)))
    return BaseIntegrals(h, ctx, n_ao, n_df, n_ri)


THREE_INDEX_FN = _lib.THREE_INDEX_FN


def setup_base_integrals_fn(X2, three_index, three_index_ri, df_sizes, n_ao, n_ri=0, context=None):
    """Function-based form (base_integrals.hpp:52-72): three_index(shell) returns the
    (n_ao(n_ao+1)/2, df_sizes[shell]) block of DF shell `shell`, three_index_ri(shell)
    the (n_ao n_ri, df_sizes[shell]) RI block (or None with n_ri = 0).  The library
    requests the shells in groups and streams them to the device."""
    ctx = context or Context.default()
    X2 = _f64(X2)
    n_df = X2.shape[0]
    sizes = [int(x) for x in df_sizes]
    P = n_ao * (n_ao + 1) // 2
    errors = []

    def wrap(fn, rows):
        def cb(_user, shell, out):
            try:
                blk = np.asarray(fn(int(shell)), dtype=np.float64)
                if blk.shape != (rows, sizes[shell]):
                    raise ValueError(f'shell {shell}: block shape {blk.shape}, expected {(rows, sizes[shell])}')
                dst = np.ctypeslib.as_array(out, shape=(rows * sizes[shell],))
                dst[:] = blk.reshape(-1, order='F')
                return 0
            except Exception as exc:  # reported through the library's error path
                errors.append(exc)
                return 1
        return THREE_INDEX_FN(cb)

    f3 = wrap(three_index, P)
    f3ri = wrap(three_index_ri, n_ao * n_ri) if three_index_ri is not None else THREE_INDEX_FN()
    sz = (ctypes.c_size_t * len(sizes))(*sizes)
    h = vp()
    rc = _lib.load().uw12g_setup_base_integrals_fn(ctx._h, _ptr(X2), f3, f3ri, None, sz, len(sizes), n_ao, n_df,
                                                   n_ri if three_index_ri is not None else 0, ctypes.byref(h))
    if rc != 0 and errors:
        raise UW12Error(rc, f'{_lib.load().uw12g_last_error().decode()}: {errors[0]}')
    check(rc)
    return BaseIntegrals(h, ctx, n_ao, n_df, n_ri if three_index_ri is not None else 0)


def setup_base_integrals_gen(X2, gen, user, n_ao, n_ri=0, stored=True, context=None, keep=None):
    """Device-generated integrals (uw12g_setup_base_integrals_gen; the function-based
    constructor with its storage flags, base_integrals.hpp:52-62).  `gen` is the address of
    a uw12g_three_index_gen_fn, `user` its user pointer; `keep` is held by the returned
    object (the generator must outlive it).  stored=False keeps only the DF metric's
    eigenvectors and regenerates the blocks at every evaluation."""
    ctx = context or Context.default()
    X2 = _f64(X2)
    n_df = X2.shape[0]
    h = vp()
    check(_lib.load().uw12g_setup_base_integrals_gen(ctx._h, _ptr(X2), vp(gen), vp(user), n_ao, n_df, n_ri,
                                                     1 if stored else 0, ctypes.byref(h)))
    b = BaseIntegrals(h, ctx, n_ao, n_df, n_ri)
    b._keep = keep
    b.stored = bool(stored)
    return b


def setup_abs_projectors(S, n_ao, n_ri, context=None):
    """libuw12.hpp:80."""
    ctx = context or Context.default()
    S = _f64(S)
    if S.shape != (n_ao + n_ri, n_ao + n_ri):
        raise UW12Error(-1, 'Total number of basis functions is inconsistent')
    h = vp()
    check(_lib.load().uw12g_setup_abs_projectors(ctx._h, _ptr(S), n_ao, n_ri, ctypes.byref(h)))
    return ABSProjectors(h, ctx, n_ao, n_ri)


def setup_orbitals(C, n_ao, n_orb, n_spin):
    """libuw12.hpp:90-92: C is n_ao x (n_orb n_spin), alpha block then beta."""
    if n_spin not in (1, 2):
        raise UW12Error(-1, 'Number of spin channels must be 1 or 2')
    C = _f64(C)
    if C.shape != (n_ao, n_orb * n_spin):
        raise UW12Error(-1, 'orbital array has inconsistent shape')
    return [np.asfortranarray(C[:, s * n_orb:(s + 1) * n_orb]) for s in range(n_spin)]


def setup_occupations(occ, n_occ_alpha, n_occ_beta=None):
    """libuw12.hpp:101 (closed shell) and :111-113 (open shell)."""
    occ = np.asarray(occ, dtype=np.float64).ravel()
    if n_occ_beta is None:
        return [occ[:n_occ_alpha].copy()]
    return [occ[:n_occ_alpha].copy(), occ[n_occ_alpha:n_occ_alpha + n_occ_beta].copy()]


def _orbital_args(orbitals, occ, n_active):
    n_spin = len(orbitals)
    if n_spin not in (1, 2):
        raise UW12Error(-1, 'invalid number of spin channels')
    if len(occ) != n_spin:
        raise UW12Error(-1, 'Different number of spin channels for orbitals and occupations')
    if len(n_active) != n_spin:
        raise UW12Error(-1, 'Different number of spin channels in n_active and orbitals')
    n_ao = orbitals[0].shape[0]
    n_orb = max(o.shape[1] for o in orbitals)
    C = np.zeros((n_ao, n_orb * n_spin), order='F')
    for s, o in enumerate(orbitals):
        if o.shape[0] != n_ao:
            raise UW12Error(-1, 'Inconsistent number of atomic orbitals')
        C[:, s * n_orb:s * n_orb + o.shape[1]] = o
        if len(occ[s]) > o.shape[1]:
            raise UW12Error(-1, 'more occupied orbitals than orbitals provided')
    occv = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in occ]))
    n_occ = (ctypes.c_size_t * n_spin)(*[len(x) for x in occ])
    n_act = (ctypes.c_size_t * n_spin)(*[int(x) for x in n_active])
    return C, n_ao, n_orb, n_spin, occv, n_occ, n_act


def form_fock(W, V, WV, abs_projectors, orbitals, occ, n_active, indirect_term=True, calculate_fock=True,
              scale_opp_spin=1.0, scale_same_spin=0.0, precision=0):
    """src/uw12.hpp:129-142: returns (dict of the six energy pieces + totals, list of Fock matrices or None)."""
    C, n_ao, n_orb, n_spin, occv, n_occ, n_act = _orbital_args(orbitals, occ, n_active)
    ctx = W.ctx
    terms = np.zeros(6)
    energy = ctypes.c_double(0.0)
    F = np.zeros((n_ao, n_ao * n_spin), order='F') if calculate_fock else None
    check(_lib.load().uw12g_form_fock(ctx._h, W._h, V._h, WV._h, abs_projectors._h, _ptr(C), n_ao, n_orb, n_spin,
                                      _ptr(occv), n_occ, n_act, 1 if indirect_term else 0, float(scale_opp_spin),
                                      float(scale_same_spin), int(precision), _ptr(F), _ptr(terms),
                                      ctypes.byref(energy)))
    out = dict(zip(TERMS, terms.tolist()))
    out['e2'] = out['e2_direct'] + out['e2_indirect']
    out['e3'] = out['e3_direct'] + out['e3_indirect']
    out['e4'] = out['e4_direct'] + out['e4_indirect']
    out['total'] = energy.value
    fock = [np.array(F[:, s * n_ao:(s + 1) * n_ao]) for s in range(n_spin)] if calculate_fock else None
    return out, fock


class Tape:
    """Handle of a forward call's intermediates (differentiation form)."""

    def __init__(self, ctx, tape_id, n_ao, n_spin):
        self.ctx, self.id, self.n_ao, self.n_spin = ctx, tape_id, n_ao, n_spin


def forward(W, V, WV, abs_projectors, orbitals, occ, n_active, indirect_term=True, scale_opp_spin=1.0,
            scale_same_spin=0.0, precision=0):
    """Differentiation form, forward half (uw12g_forward): returns (energy pieces as form_fock, Tape)."""
    C, n_ao, n_orb, n_spin, occv, n_occ, n_act = _orbital_args(orbitals, occ, n_active)
    ctx = W.ctx
    terms = np.zeros(6)
    energy = ctypes.c_double(0.0)
    tape = ctypes.c_ulonglong(0)
    check(_lib.load().uw12g_forward(ctx._h, W._h, V._h, WV._h, abs_projectors._h, _ptr(C), n_ao, n_orb, n_spin,
                                    _ptr(occv), n_occ, n_act, 1 if indirect_term else 0, float(scale_opp_spin),
                                    float(scale_same_spin), int(precision), _ptr(terms), ctypes.byref(energy),
                                    ctypes.byref(tape)))
    out = dict(zip(TERMS, terms.tolist()))
    out['total'] = energy.value
    return out, Tape(ctx, tape.value, n_ao, n_spin)


def adjoint(tape):
    """Differentiation form, adjoint half (uw12g_adjoint): F = dE/dD per spin from the forward's intermediates."""
    F = np.zeros((tape.n_ao, tape.n_ao * tape.n_spin), order='F')
    check(_lib.load().uw12g_adjoint(tape.ctx._h, ctypes.c_ulonglong(tape.id), _ptr(F)))
    return [np.array(F[:, s * tape.n_ao:(s + 1) * tape.n_ao]) for s in range(tape.n_spin)]


def uw12_energy(W, V, WV, abs_projectors, orbitals, occ, n_active, scale_opp_spin, scale_same_spin,
                print_level=0):
    """libuw12.hpp:129-140."""
    C, n_ao, n_orb, n_spin, occv, n_occ, n_act = _orbital_args(orbitals, occ, n_active)
    e = ctypes.c_double(0.0)
    check(_lib.load().uw12g_energy(W.ctx._h, W._h, V._h, WV._h, abs_projectors._h, _ptr(C), n_ao, n_orb, n_spin,
                                   _ptr(occv), n_occ, n_act, float(scale_opp_spin), float(scale_same_spin),
                                   ctypes.byref(e)))
    if print_level > 0:
        print(f'Total UW12 energy {e.value:20.12f}')
    return e.value


def uw12_fock(W, V, WV, abs_projectors, orbitals, occ, n_active, scale_opp_spin, scale_same_spin,
              print_level=0):
    """libuw12.hpp:157-169: returns (energy, fock) with fock n_ao x (n_ao n_spin), alpha block then beta."""
    C, n_ao, n_orb, n_spin, occv, n_occ, n_act = _orbital_args(orbitals, occ, n_active)
    e = ctypes.c_double(0.0)
    F = np.zeros((n_ao, n_ao * n_spin), order='F')
    check(_lib.load().uw12g_fock(W.ctx._h, _ptr(F), W._h, V._h, WV._h, abs_projectors._h, _ptr(C), n_ao, n_orb,
                                 n_spin, _ptr(occv), n_occ, n_act, float(scale_opp_spin), float(scale_same_spin),
                                 ctypes.byref(e)))
    if print_level > 0:
        print(f'Total UW12 energy {e.value:20.12f}')
    return e.value, F
