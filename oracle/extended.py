"""UW12 energy terms in extended precision, for small systems.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this file;
only ``tests/`` and ``bench.py``'s CPU-baseline leg may use anything under
``oracle/``, and only as a checker.

Why it exists: on ill-conditioned inputs (the s-Gaussian recipe keeps DF-metric
eigenvalues down to ~2.5e-10 and ABS-overlap eigenvalues down to ~1.1e-6) two
FP64 evaluations of the same formulas that differ only in association already
disagree by ~1e-9 Eh, so neither can serve as the exact value the other is
checked against.  This module evaluates the same quantities with every product
and sum in ``numpy.longdouble`` (x86 80-bit, 64-bit significand: 2^11 times
finer than FP64), from the FP64 eigendecompositions the reference itself uses,
so its own rounding sits ~1e-13 Eh or below at these sizes and what remains
between an FP64 result and this one is that FP64 result's rounding.

Formulas (the documented index sums, as in ``brute.py``):

* DF four-index integrals in the metric eigenbasis:
  (pq|X|rs) = sum_A (pq|A~) (1/lambda_A) (A~|rs), (pq|A~) = sum_B (pq|B) Q_BA,
  keeping |lambda| > 1e-10 -- src/integrals/base_integrals.cpp:155-166,362-373
  (P2 = Q pinv(Lambda), J3 = X3 P2, weights lambda), src/utils/linalg.hpp:469-481.
* ABS+ projector S+ = sum over |sigma| > max(1e-6, 1e-8 max|sigma|) of
  v v^T / sigma -- src/three_electron/ri_utils.hpp:45-74,
  src/utils/linalg.hpp:1157-1197.
* E2 direct/indirect -- src/two_electron/two_electron.cpp:58-69,101-114,148-150.
* E3 direct -- src/three_electron/direct_utils.cpp:136-157 (sign :152);
  E3 indirect -- src/three_electron/indirect_utils.cpp:19-133 (sign :124),
  x kappa1 src/three_electron/three_electron.cpp:183-186.
* E4 direct -- src/four_electron/four_electron.cpp:118-119; E4 indirect
  -- :143-179,307-309.
* spin factors -- src/four_electron/four_electron_utils.cpp:9-30.

Only the eigendecompositions run in FP64 (``numpy.linalg.eigh``, LAPACK
syevd, the routine behind the reference's ``eig_sym``).  Sizes must stay tiny
(n_ao <= ~48): the products run without BLAS.
"""
import numpy as np

from .brute import spin_factor, unpack_lower

LD = np.longdouble


def _es(*args):
    return np.einsum(*args, optimize=True)


def _df(X3, X2, X3ri, N):
    """(AO-pair|A~) as [mu, nu, A], (RI-AO|A~) as [r, mu, A] (or None), weights 1/lambda."""
    lam, Q = np.linalg.eigh(0.5 * (X2 + X2.T))
    keep = np.abs(lam) > 1e-10
    Qk = Q[:, keep].astype(LD)
    w = (1 / lam[keep].astype(LD))
    T = _es('mnB,BA->mnA', unpack_lower(X3, N).astype(LD), Qk)
    if X3ri is None:
        return T, None, w
    Nd = X3.shape[1]
    Nr = X3ri.shape[0] // N
    Tri = X3ri.reshape(N, Nr, Nd).transpose(1, 0, 2).astype(LD)  # row r + Nr mu -> [r, mu, B]
    return T, _es('rmB,BA->rmA', Tri, Qk), w


def abs_pinv(S):
    """S+ from the FP64 eigendecomposition of S, formed in extended precision."""
    sig, V = np.linalg.eigh(0.5 * (S + S.T))
    thr = max(np.abs(sig).max() * 1e-8, 1e-6)
    keep = np.abs(sig) > thr
    Vk = V[:, keep].astype(LD)
    return _es('pm,m,qm->pq', Vk, 1 / sig[keep].astype(LD), Vk)


def energy_terms(ints, S, C, occ, n_active, k0, k1, indirect=True, raw=False):
    """Six energy pieces and their total as Python floats (rounded from long double),
    or as numpy.longdouble values with raw=True.

    Same arguments as ``brute.energy_terms``: ints {'W', 'V', 'WV': (X3, X2, X3ri)},
    C / occ / n_active per spin; C may be given in long double (finite differences
    along orbital paths then carry no FP64 rounding of the perturbed orbitals)."""
    n_spin = len(C)
    N = C[0].shape[0]
    g = 2.0 if n_spin == 1 else 1.0
    e = g
    Co = [np.asarray(C[s]).astype(LD)[:, :len(occ[s])] * np.sqrt(np.asarray(occ[s], dtype=LD) / LD(g))
          for s in range(n_spin)]
    Ca = [Co[s][:, Co[s].shape[1] - n_active[s]:] for s in range(n_spin)]
    do_ind = indirect and k1 != 0.0
    out = {}

    # two-electron (WV, active orbitals)
    T, _, w = _df(*ints['WV'], N)
    aa = [_es('mi,mnA,nj->ijA', Ca[s], T, Ca[s]) for s in range(n_spin)]
    d = [_es('iiA->A', aa[s]) for s in range(n_spin)]
    e2d = LD(0)
    for s in range(n_spin):
        for t in range(n_spin):
            e2d += LD(0.5 * spin_factor(n_spin, s, t, k0, k1)) * np.sum(d[s] * w * d[t])
    e2i = LD(0)
    if do_ind:
        for s in range(n_spin):
            e2i -= LD(0.5 * k1 * e) * _es('ijA,A,ijA->', aa[s], w, aa[s])
    out['e2_direct'], out['e2_indirect'] = e2d, e2i

    TW, RW, wW = _df(*ints['W'], N)
    TV, RV, wV = _df(*ints['V'], N)
    GW = np.concatenate([TW, RW], axis=0)  # [mu', nu, A], mu' over AO then RI
    GV = np.concatenate([TV, RV], axis=0)
    Sp = abs_pinv(S)

    def mo(T, Cl, Cr):
        return _es('mi,mnA,nk->ikA', Cl, T, Cr)

    e3d = e4d = e3i = e4i = LD(0)
    for s in range(n_spin):
        for t in range(n_spin):
            f = LD(spin_factor(n_spin, s, t, k0, k1))
            if f == 0:
                continue
            Wik = mo(TW, Ca[t], Co[t])
            Wjl = mo(TW, Ca[s], Co[s])
            Vik = mo(TV, Ca[t], Co[t])
            Vjl = mo(TV, Ca[s], Co[s])
            tt = _es('ikA,ikB->AB', Wik, Vik)  # sum_ik (ik|A~)_W (ik|B~)_V, spin t
            ts = _es('jlA,jlB->AB', Wjl, Vjl)  # spin s
            # E4 direct: 1/2 f sum (ik|W|jl)(ik|V|jl) = 1/2 f sum_AB wW_A wV_B tt[A,B] ts[A,B]
            e4d += LD(0.5) * f * _es('A,AB,AB,B->', wW, tt, ts, wV)
            # E3 direct: -f sum (j mu'|W|ik) S+ (nu' j|V|ik), j act in s; i act, k occ in t
            PW = _es('pnA,nj->pjA', GW, Ca[s])
            PV = _es('pnA,nj->pjA', GV, Ca[s])
            X = _es('pjA,pq,qjB->AB', PW, Sp, PV)
            e3d -= f * _es('A,AB,AB,B->', wW, X, tt, wV)
        if do_ind:
            Wx = mo(TW, Ca[s], Co[s])  # (ik|A~), i act, k occ
            Vx = mo(TV, Ca[s], Co[s])
            W4 = _es('ilA,A,jkA->iljk', Wx, wW, Wx)
            V4 = _es('ikA,A,jlA->ikjl', Vx, wV, Vx)
            e4i -= LD(0.5 * k1 * e) * _es('iljk,ikjl->', W4, V4)
            PWi = _es('pnA,ni->piA', GW, Ca[s])
            PVj = _es('pnA,nj->pjA', GV, Ca[s])
            W3 = _es('piA,A,jkA->pijk', PWi, wW, Wx)  # (mu' i|W|jk)
            V3 = _es('qjA,A,ikA->qjik', PVj, wV, Vx)  # (nu' j|V|ik)
            e3i += LD(k1 * e) * _es('pijk,pq,qjik->', W3, Sp, V3)
    out['e3_direct'], out['e3_indirect'] = e3d, e3i
    out['e4_direct'], out['e4_indirect'] = e4d, e4i
    out['total'] = e2d + e2i + e3d + e3i + e4d + e4i
    if raw:
        return out
    return {k: float(v) for k, v in out.items()}
