"""Brute-force UW12 energy terms from explicit four-index DF integrals.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this file;
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may use anything under ``oracle/``, and only as a checker.

This module pins the *formulas* of the UW12 energy, independently of both the
reference's loop order and the GPU path's GEMM re-association: it builds the
density-fitted four-index integrals

    (pq|X|rs) = sum_AB (pq|A)_X [X2^+]_AB (B|rs)_X

explicitly as dense arrays and evaluates each energy term as the plain index
sum written in the reference's documentation:

* ``E2`` -- src/uw12.hpp:33-35 (definition), src/two_electron/two_electron.hpp:11-43,
  spin factors src/two_electron/two_electron.cpp:58-69, indirect sign/scale
  two_electron.cpp:101-114,148-150.
* ``E3`` -- src/uw12.hpp:36-37, src/three_electron/three_electron.hpp:16-155;
  direct sign direct_utils.cpp:152, indirect sign indirect_utils.cpp:124 and
  three_electron.cpp:183-186.
* ``E4`` -- src/uw12.hpp:38-40, src/four_electron/four_electron.hpp:14-63;
  direct four_electron.cpp:118-119, indirect four_electron.cpp:160-178,307-309.
* spin factor f(s,s') -- src/four_electron/four_electron_utils.cpp:9-30.
* DF metric pseudo-inverse (absolute 1e-10 on |lambda|) --
  src/integrals/base_integrals.cpp:155-166, src/utils/linalg.hpp:469-481.
* ABS+ projector S^+ (drop |sigma| <= max(1e-6, 1e-8 max|sigma|)) --
  src/three_electron/ri_utils.hpp:45-74, src/utils/linalg.hpp:1157-1197.

Sizes must stay tiny (n_ao <= ~24): every four-index array is built densely.
"""
import numpy as np


def _es(*args):
    return np.einsum(*args, optimize=True)


def unpack_lower(packed, n):
    """(n(n+1)/2, k) packed-lower rows (idx = i(i+1)/2 + j, j <= i;
    src/utils/utils.hpp:23-48) -> (n, n, k) symmetric."""
    k = packed.shape[1]
    out = np.zeros((n, n, k))
    il, jl = np.tril_indices(n)
    # np.tril_indices enumerates (i, j) row by row with j <= i: i(i+1)/2 + j
    out[il, jl, :] = packed
    out[jl, il, :] = packed
    return out


def metric_pinv(X2, tol=1e-10):
    vals, vecs = np.linalg.eigh(0.5 * (X2 + X2.T))
    inv = np.where(np.abs(vals) > tol, 1.0 / np.where(vals == 0, 1, vals), 0.0)
    return (vecs * inv) @ vecs.T


def abs_pinv(S, rel=1e-8, abs_=1e-6):
    vals, vecs = np.linalg.eigh(0.5 * (S + S.T))
    thr = max(np.abs(vals).max() * rel, abs_)
    keep = np.abs(vals) > thr
    v = vecs[:, keep]
    return (v / vals[keep]) @ v.T


def spin_factor(n_spin, s, sp, k0, k1):
    if n_spin == 1:
        return 2.0 * (k0 + k1)
    return k1 if s == sp else k0


def weighted_orbitals(C, occ, n_spin):
    g = 2.0 if n_spin == 1 else 1.0
    return [C[s][:, :len(occ[s])] * np.sqrt(np.asarray(occ[s]) / g) for s in range(n_spin)]


def energy_terms(ints, S, C, occ, n_active, k0, k1, indirect=True):
    """Return dict of UW12 energy pieces.

    ints: {'W': (X3, X2, X3ri), 'V': (X3, X2, X3ri), 'WV': (X3, X2, None)} with
    X3 packed (n_ao(n_ao+1)/2, n_df), X3ri (n_ao*n_ri, n_df) with row r + n_ri*mu.
    C: list per spin of (n_ao, n_orb); occ: list per spin; n_active: list.
    """
    n_spin = len(C)
    N = C[0].shape[0]
    e = 2.0 if n_spin == 1 else 1.0
    Co = weighted_orbitals(C, occ, n_spin)
    Ca = [Co[s][:, Co[s].shape[1] - n_active[s]:] for s in range(n_spin)]
    do_ind = indirect and k1 != 0.0

    def ao3(name):
        X3, X2, X3ri = ints[name]
        T = unpack_lower(X3, N)
        if X3ri is None:
            return T, None, metric_pinv(X2)
        Nd = X3.shape[1]
        Nr = X3ri.shape[0] // N
        Tri = X3ri.reshape(N, Nr, Nd).transpose(1, 0, 2)  # [r, mu, A]
        G = np.concatenate([T, Tri], axis=0)  # [mu', nu, A], mu' over AO then RI
        return T, G, metric_pinv(X2)

    out = {}
    # ---------------- two-electron (WV, active orbitals only) ----------------
    T, _, Minv = ao3('WV')
    aa = [_es('mi,mnA,nj->ijA', Ca[s], T, Ca[s]) for s in range(n_spin)]
    I4 = [[_es('ijA,AB,klB->ijkl', aa[s], Minv, aa[t]) for t in range(n_spin)]
          for s in range(n_spin)]
    e2d = 0.0
    for s in range(n_spin):
        for t in range(n_spin):
            f = spin_factor(n_spin, s, t, k0, k1)
            e2d += 0.5 * f * _es('iijj->', I4[s][t])
    e2i = 0.0
    if do_ind:
        for s in range(n_spin):
            e2i -= 0.5 * k1 * e * _es('ijij->', I4[s][s])
    out['e2_direct'], out['e2_indirect'] = e2d, e2i

    # ---------------- W and V four-index pieces ----------------
    TW, GW, MW = ao3('W')
    TV, GV, MV = ao3('V')
    Sp = abs_pinv(S)
    # (ik|X|A) with i active, k occupied; (mu' i|X|A) with i active
    def mo(T, Cl, Cr):
        return _es('mi,mnA,nk->ikA', Cl, T, Cr)

    e4d = 0.0
    e3d = 0.0
    e4i = 0.0
    e3i = 0.0
    for s in range(n_spin):
        for t in range(n_spin):
            f = spin_factor(n_spin, s, t, k0, k1)
            # E4 direct: sum (ik|W|jl)(ik|V|jl), i act/k occ in t, j act/l occ in s
            Wik = mo(TW, Ca[t], Co[t]); Wjl = mo(TW, Ca[s], Co[s])
            Vik = mo(TV, Ca[t], Co[t]); Vjl = mo(TV, Ca[s], Co[s])
            W4 = _es('ikA,AB,jlB->ikjl', Wik, MW, Wjl)
            V4 = _es('ikA,AB,jlB->ikjl', Vik, MV, Vjl)
            e4d += 0.5 * f * np.sum(W4 * V4)
            # E3 direct: -f sum (j mu'|W|ik) S+ (nu' j|V|ik), j act in s; i act, k occ in t
            GWj = _es('pnA,nj->pjA', GW, Ca[s])
            GVj = _es('pnA,nj->pjA', GV, Ca[s])
            W3 = _es('pjA,AB,ikB->pjik', GWj, MW, Wik)
            V3 = _es('qjA,AB,ikB->qjik', GVj, MV, Vik)
            e3d -= f * _es('pjik,pq,qjik->', W3, Sp, V3)
        if do_ind:
            Wx = mo(TW, Ca[s], Co[s]); Vx = mo(TV, Ca[s], Co[s])
            W4 = _es('ilA,AB,jkB->iljk', Wx, MW, Wx)
            V4 = _es('ikA,AB,jlB->ikjl', Vx, MV, Vx)
            e4i -= 0.5 * k1 * e * _es('iljk,ikjl->', W4, V4)
            GWi = _es('pnA,ni->piA', GW, Ca[s])
            GVj = _es('pnA,nj->pjA', GV, Ca[s])
            W3 = _es('piA,AB,jkB->pijk', GWi, MW, Wx)   # (mu' i|W|jk)
            V3 = _es('qjA,AB,ikB->qjik', GVj, MV, Vx)   # (nu' j|V|ik)
            e3i += k1 * e * _es('pijk,pq,qjik->', W3, Sp, V3)
    out['e3_direct'], out['e3_indirect'] = e3d, e3i
    out['e4_direct'], out['e4_indirect'] = e4d, e4i
    out['e2'] = e2d + e2i
    out['e3'] = e3d + e3i
    out['e4'] = e4d + e4i
    out['total'] = out['e2'] + out['e3'] + out['e4']
    return out
