"""Synthetic UW12 inputs: s-Gaussian clusters (the benchmark configurations)
and small well-conditioned random systems (the style of the reference's own
unit-test fixtures, tests/setup_integrals.hpp, src/utils/linalg.hpp:277-292,507-516).

The integrals are computed by libuw12_synth.so (csrc/synth.cpp, host OpenMP);
geometry, bases and orbitals are built here with numpy's PCG64 generator.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'libuw12_synth.so')
DEV_LIB = os.path.join(HERE, 'libuw12_synth_dev.so')
_lib = None
_dev_lib = None

# per-atom basis recipe: (count, first exponent, ratio)
AO = (8, 0.1, 2.5)
DF = (24, 0.2, 2.0)
RI = (32, 0.15, 1.8)
GEMINAL = ((0.4, 0.6), (0.5, 2.0))  # W = -sum c_i exp(-g_i r^2)

# name -> (atoms, box, min separation, (n_alpha, n_beta or None), truncate last atom to (ao, df, ri))
CONFIGS = {
    'c1': dict(atoms=3, box=4.0, min_sep=2.0, occ=(5, None)),
    'c2': dict(atoms=40, box=20.0, min_sep=2.5, occ=(80, None)),
    'c3': dict(atoms=100, box=30.0, min_sep=2.5, occ=(160, 150)),
    'c4': dict(atoms=100, box=30.0, min_sep=2.5, occ=(160, None)),
    'c5': dict(atoms=188, box=40.0, min_sep=2.5, occ=(300, 290), last=(4, 12, 16)),
    # intermediate sizes for parity tests (same recipe and atom density as c2)
    'm6': dict(atoms=6, box=10.6, min_sep=2.5, occ=(12, None)),
    'm6u': dict(atoms=6, box=10.6, min_sep=2.5, occ=(12, 11)),
    'm12': dict(atoms=12, box=13.4, min_sep=2.5, occ=(24, None)),
    'm12u': dict(atoms=12, box=13.4, min_sep=2.5, occ=(24, 23)),
}
CONFIG_INDEX = {'c1': 1, 'c2': 2, 'c3': 3, 'c4': 4, 'c5': 5, 'm6': 6, 'm6u': 7, 'm12': 8, 'm12u': 9}


def build():
    subprocess.check_call(['g++', '-O2', '-std=c++17', '-fopenmp', '-fPIC', '-shared', '-o', LIB,
                           os.path.join(HERE, 'csrc', 'synth.cpp')])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.synth_set_geminal.argtypes = [ctypes.c_int, dp, dp]
        L.synth_overlap.argtypes = [dp, dp, ctypes.c_int, dp]
        L.synth_two_index.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int, dp]
        L.synth_three_index.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, ctypes.c_int, dp]
        L.synth_three_index_ri.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp,
                                           ctypes.c_int, dp]
        L.synth_pair_integral.argtypes = [ctypes.c_int, ctypes.c_double, dp, ctypes.c_double, dp]
        L.synth_pair_integral.restype = ctypes.c_double
        L.synth_bare.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.synth_bare.restype = ctypes.c_double
        c = np.array(GEMINAL[0], dtype=np.float64)
        g = np.array(GEMINAL[1], dtype=np.float64)
        L.synth_set_geminal(len(c), c.ctypes.data_as(dp), g.ctypes.data_as(dp))
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


KINDS = {'V': 0, 'W': 1, 'WV': 2}


def bare_integral(kind, p, q, R2, gamma=0.0):
    """Two-centre integral of unit Gaussian distributions exp(-p r^2), exp(-q r^2) at
    squared distance R2 under 1/r (kind 'V'), exp(-g r^2) ('W'), exp(-g r^2)/r ('WV')."""
    return lib().synth_bare(KINDS[kind], float(p), float(q), float(R2), float(gamma))


def geometry(n_atoms, box, min_sep, seed):
    """Uniform positions in a cube with a minimum separation (rejection sampling)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = []
    tries = 0
    while len(pts) < n_atoms:
        p = rng.uniform(0.0, box, 3)
        tries += 1
        if tries > 10_000_000:
            raise RuntimeError('could not place atoms')
        if all(np.sum((p - q) ** 2) >= min_sep ** 2 for q in pts):
            pts.append(p)
    return np.array(pts)


def basis(atoms, recipe, last=None):
    n, a0, r = recipe
    ex = a0 * r ** np.arange(n)
    exps, cens = [], []
    for i, at in enumerate(atoms):
        e = ex if (last is None or i < len(atoms) - 1) else ex[:last]  # most diffuse first
        exps.append(e)
        cens.append(np.repeat(at[None, :], len(e), axis=0))
    return np.ascontiguousarray(np.concatenate(exps)), np.ascontiguousarray(np.concatenate(cens))


def overlap(ex, cen):
    n = len(ex)
    S = np.zeros((n, n), order='F')
    lib().synth_overlap(_p(ex), _p(cen), n, _p(S))
    return 0.5 * (S + S.T)


def two_index(kind, ex, cen):
    n = len(ex)
    X = np.zeros((n, n), order='F')
    lib().synth_two_index(KINDS[kind], _p(ex), _p(cen), n, _p(X))
    return np.asfortranarray(0.5 * (X + X.T))


def three_index(kind, eao, cao, edf, cdf):
    N, Nd = len(eao), len(edf)
    X = np.zeros((N * (N + 1) // 2, Nd), order='F')
    lib().synth_three_index(KINDS[kind], _p(eao), _p(cao), N, _p(edf), _p(cdf), Nd, _p(X))
    return X


def three_index_ri(kind, eri, cri, eao, cao, edf, cdf):
    N, Nr, Nd = len(eao), len(eri), len(edf)
    X = np.zeros((N * Nr, Nd), order='F')
    lib().synth_three_index_ri(KINDS[kind], _p(eri), _p(cri), Nr, _p(eao), _p(cao), N, _p(edf), _p(cdf), Nd,
                               _p(X))
    return X


def lowdin_orbitals(S_ao, n_occ, rng):
    G = rng.standard_normal((S_ao.shape[0], n_occ))
    w, U = np.linalg.eigh(G.T @ S_ao @ G)
    return np.asfortranarray(G @ (U / np.sqrt(w)) @ U.T)


def make_system(name, with_integrals=True, seed_offset=0):
    """Inputs of a named benchmark configuration as a dict with keys
    N, Nd, Nr, S, ints {'W','V','WV': (X3, X2, X3_ri)}, orbitals, occ, n_active, atoms."""
    cfg = CONFIGS[name]
    base = 1000 * CONFIG_INDEX[name] + seed_offset
    atoms = geometry(cfg['atoms'], cfg['box'], cfg['min_sep'], base + 0)
    last = cfg.get('last')
    eao, cao = basis(atoms, AO, last[0] if last else None)
    edf, cdf = basis(atoms, DF, last[1] if last else None)
    eri, cri = basis(atoms, RI, last[2] if last else None)
    N, Nd, Nr = len(eao), len(edf), len(eri)
    sys_ = dict(name=name, N=N, Nd=Nd, Nr=Nr, atoms=atoms, basis=dict(ao=(eao, cao), df=(edf, cdf), ri=(eri, cri)))
    S = overlap(np.concatenate([eao, eri]), np.concatenate([cao, cri]))
    sys_['S'] = S
    rng = np.random.Generator(np.random.PCG64(base + 1))
    na, nb = cfg['occ']
    Ca = lowdin_orbitals(S[:N, :N], na, rng)
    if nb is None:
        sys_['orbitals'] = [Ca]
        sys_['occ'] = [np.full(na, 2.0)]
        sys_['n_active'] = [na]
    else:
        rng2 = np.random.Generator(np.random.PCG64(base + 2))
        R, _ = np.linalg.qr(rng2.standard_normal((na, na)))
        Cb = np.asfortranarray(Ca @ R)
        sys_['orbitals'] = [Ca, Cb]
        sys_['occ'] = [np.full(na, 1.0), np.full(nb, 1.0)]
        sys_['n_active'] = [na, nb]
    if with_integrals:
        ints = {}
        for kind in ('W', 'V', 'WV'):
            X2 = two_index(kind, edf, cdf)
            X3 = three_index(kind, eao, cao, edf, cdf)
            X3ri = three_index_ri(kind, eri, cri, eao, cao, edf, cdf) if kind != 'WV' else None
            ints[kind] = (X3, X2, X3ri)
        sys_['ints'] = ints
    return sys_


def random_system(n_ao, n_df, n_ri, n_occ, n_active=None, seed=11, uniform_orbitals=True):
    """Well-conditioned random inputs in the style of the reference's unit tests:
    uniform [0,1) three-index arrays, metrics U diag(1 + u) U^T with U = expm(skew)
    (condition number <= 2; W and WV negative definite like a geminal), S positive definite."""
    from scipy.linalg import expm
    rng = np.random.Generator(np.random.PCG64(seed))
    P = n_ao * (n_ao + 1) // 2

    def rpd(n, sign=1.0):
        X = rng.uniform(size=(n, n))
        U = expm(0.5 * (X - X.T))
        ev = 1.0 + rng.uniform(size=n)
        M = sign * (U * ev) @ U.T
        return np.asfortranarray(0.5 * (M + M.T))

    ints = {}
    for kind, sign in (('W', -1.0), ('V', 1.0), ('WV', -1.0)):
        X3 = np.asfortranarray(rng.uniform(size=(P, n_df)))
        X3ri = np.asfortranarray(rng.uniform(size=(n_ao * n_ri, n_df))) if kind != 'WV' else None
        ints[kind] = (X3, rpd(n_df, sign), X3ri)
    S = rpd(n_ao + n_ri)
    n_spin = len(n_occ)
    n_orb = max(n_occ) + 1
    orbitals = [np.asfortranarray(rng.uniform(size=(n_ao, n_orb)) if uniform_orbitals
                                  else rng.standard_normal((n_ao, n_orb))) for _ in range(n_spin)]
    occ = [rng.uniform(0.5, 1.0, size=o) * (2.0 if n_spin == 1 else 1.0) for o in n_occ]
    if n_active is None:
        n_active = list(n_occ)
    return dict(name='random', N=n_ao, Nd=n_df, Nr=n_ri, S=S, ints=ints, orbitals=orbitals, occ=occ,
                n_active=list(n_active))


def shell_integrals(system, kind):
    """Per-DF-shell generators for one operator of a make_system() system, the
    reference's ThreeIndexFn form (integral_functions.hpp:16-55): a shell is the
    set of DF functions on one atom.  Returns (X2, df_sizes, fn, fn_ri) where
    fn(shell) / fn_ri(shell) compute that shell's packed-AO / RI columns on
    demand (fn_ri is None for WV)."""
    eao, cao = system['basis']['ao']
    edf, cdf = system['basis']['df']
    eri, cri = system['basis']['ri']
    starts = [0] + [i for i in range(1, len(edf)) if not np.array_equal(cdf[i], cdf[i - 1])] + [len(edf)]
    sizes = [b - a for a, b in zip(starts[:-1], starts[1:])]

    def fn(shell):
        a, b = starts[shell], starts[shell + 1]
        return three_index(kind, eao, cao, np.ascontiguousarray(edf[a:b]), np.ascontiguousarray(cdf[a:b]))

    def fn_ri(shell):
        a, b = starts[shell], starts[shell + 1]
        return three_index_ri(kind, eri, cri, eao, cao, np.ascontiguousarray(edf[a:b]),
                              np.ascontiguousarray(cdf[a:b]))

    return two_index(kind, edf, cdf), sizes, fn, (fn_ri if kind != 'WV' else None)


def dev_lib():
    """libuw12_synth_dev.so: the device generator (csrc/synth_dev.hip), built by
    libuw12_amd.build.build_synth_dev()."""
    global _dev_lib
    if _dev_lib is None:
        if not os.path.exists(DEV_LIB):
            raise RuntimeError(f'{DEV_LIB} is not built; run `python -m libuw12_amd.build`')
        L = ctypes.CDLL(DEV_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.uw12s_create.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp,
                                   ctypes.c_int, ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_void_p)]
        L.uw12s_destroy.argtypes = [ctypes.c_void_p]
        L.uw12s_destroy.restype = None
        L.uw12s_generate.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.uw12s_generate_host.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, dp]
        _dev_lib = L
    return _dev_lib


class DeviceGenerator:
    """Three-index integrals of one operator of a make_system() system, produced on
    the device block by block (the uw12g_three_index_gen_fn callback of
    uw12g_setup_base_integrals_gen).  `fn` is the callback's address and `user`
    its user pointer; the object must outlive every integrals handle built on it."""

    def __init__(self, system, kind, device=0):
        L = dev_lib()
        eao, cao = system['basis']['ao']
        edf, cdf = system['basis']['df']
        eri, cri = system['basis']['ri']
        if kind == 'WV':
            eri, cri = np.zeros(0), np.zeros((0, 3))
        c = np.array(GEMINAL[0], dtype=np.float64)
        g = np.array(GEMINAL[1], dtype=np.float64)
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (eao, cao, edf, cdf, eri, cri)]
        self._arrs = arrs
        h = ctypes.c_void_p()
        rc = L.uw12s_create(int(device), KINDS[kind], _p(arrs[0]), _p(arrs[1]), len(eao), _p(arrs[2]), _p(arrs[3]),
                            len(edf), _p(arrs[4]) if len(eri) else None, _p(arrs[5]) if len(eri) else None,
                            len(eri), len(c), _p(c), _p(g), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f'uw12s_create failed ({rc})')
        self._h = h
        self.user = h.value
        self.fn = ctypes.cast(L.uw12s_generate, ctypes.c_void_p).value
        self.n_ao, self.n_df, self.n_ri = len(eao), len(edf), len(eri)

    def block(self, df0, n, with_ri):
        """Host copy [mu', nu, j] of the block the callback writes for DF functions [df0, df0 + n)."""
        rows = self.n_ao + (self.n_ri if with_ri else 0)
        out = np.zeros(rows * self.n_ao * n)
        rc = dev_lib().uw12s_generate_host(self._h, df0, n, int(bool(with_ri)), _p(out))
        if rc != 0:
            raise RuntimeError(f'uw12s_generate_host failed ({rc})')
        return out.reshape(n, self.n_ao, rows).transpose(2, 1, 0)

    def close(self):
        if getattr(self, '_h', None):
            dev_lib().uw12s_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_integrals(system, ctx, stored=True, device=0):
    """W, V and WV integrals of a make_system() system (built with or without host
    integrals) through the device generator: (W, V, WV) handles, each holding its
    generator.  stored=False gives the non-stored (regenerating) form for W and V;
    WV follows the same choice."""
    import libuw12_amd as u
    edf, cdf = system['basis']['df']
    out = []
    for kind in ('W', 'V', 'WV'):
        gen = DeviceGenerator(system, kind, device)
        X2 = two_index(kind, edf, cdf)
        out.append(u.setup_base_integrals_gen(X2, gen.fn, gen.user, system['N'], gen.n_ri, stored=stored,
                                              context=ctx, keep=gen))
    return tuple(out)
