"""The repeated-evaluation (SCF) pattern of the reference's example (example/example.cpp:109-219):
integrals and ABS projectors set up once, the energy and Fock evaluated again and again as the
orbitals change.  The library keeps device state on the integrals between evaluations (the
rotated Gamma and, in the split forms, its plane tensors or row maxima); every evaluation on that
kept state must give the same bits as one on freshly made handles, also when the projector
changes between evaluations."""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error, gpu_setup
from libuw12_amd import synth

pytestmark = pytest.mark.gpu


def moved_orbitals(C, S, rng, t):
    """C + t R made S-orthonormal again (Loewdin): the occupied space moves, as in an SCF step."""
    X = C + t * rng.standard_normal(C.shape)
    M = X.T @ S @ X
    w, Q = np.linalg.eigh(M)
    return X @ (Q * w ** -0.5) @ Q.T


def perturbed_overlap(S, rng, eps):
    """A second positive-definite ABS overlap near S (a different projector)."""
    Z = rng.standard_normal(S.shape)
    return S + eps * (Z @ Z.T) / S.shape[0]


@pytest.mark.parametrize('form', [u.FP64, u.MIXED_FP16X3, u.MIXED_BF16X6])
def test_scf_loop_on_kept_state_matches_fresh_handles(form):
    print(f'case m12u SCF loop form {form}', flush=True)
    s = synth.make_system('m12u')
    rng = np.random.default_rng(7)
    N = s['N']
    S2 = perturbed_overlap(s['S'], rng, 0.02)
    steps = []
    C = [np.array(c) for c in s['orbitals']]
    for it in range(4):
        C = [moved_orbitals(c, s['S'][:N, :N], rng, 0.05) for c in C]
        steps.append(([c.copy() for c in C], s['S'] if it != 2 else S2))
    kw = dict(indirect_term=True, scale_opp_spin=1.0, scale_same_spin=0.5, precision=form)
    ctx = u.Context(0)
    W, V, WV, P = gpu_setup(s, ctx)
    P2 = u.setup_abs_projectors(S2, N, s['Nr'], context=ctx)
    last = None
    for it, (orb, Sx) in enumerate(steps):
        Px = P if Sx is s['S'] else P2
        e, F = u.form_fock(W, V, WV, Px, orb, s['occ'], s['n_active'], **kw)
        fctx = u.Context(0)
        fs = dict(s, S=Sx)
        fW, fV, fWV, fP = gpu_setup(fs, fctx)
        ef, Ff = u.form_fock(fW, fV, fWV, fP, orb, s['occ'], s['n_active'], **kw)
        for k in TERMS + ('total',):
            assert e[k] == ef[k], (it, k, e[k], ef[k])
        for a, b in zip(F, Ff):
            assert np.array_equal(a, b), (it, np.abs(a - b).max())
        if last is not None:  # the orbitals moved: so did the energy
            assert e['total'] != last
        last = e['total']
        del fW, fV, fWV, fP
        fctx.close()
    # the reduced forms stay within the gate of FP64 on the last step
    if form != u.FP64:
        orb, Sx = steps[-1]
        e64, F64 = u.form_fock(W, V, WV, P, orb, s['occ'], s['n_active'], **dict(kw, precision=u.FP64))
        for k in TERMS:
            assert abs(e[k] - e64[k]) <= 1e-6, (k, e[k], e64[k])
        assert fock_error(F, F64) <= 1e-6
    del W, V, WV, P, P2
    ctx.close()


@pytest.mark.parametrize('form', [u.MIXED_FP16X3, u.MIXED_BF16X3])
def test_kept_gamma_modes_give_the_same_bits(form):
    """The state kept for the t~ dX product (the rotated V integrals' plane tensors, or for the FP16
    form only their row maxima when the planes would not fit) gives the same bits as splitting them
    within every product, over two evaluations with moving orbitals."""
    from libuw12_amd import _lib
    L = _lib.load()
    s = synth.make_system('m12u')
    rng = np.random.default_rng(11)
    N = s['N']
    kw = dict(indirect_term=True, scale_opp_spin=1.0, scale_same_spin=0.5, precision=form)
    orbs = [[np.array(c) for c in s['orbitals']]]
    orbs.append([moved_orbitals(c, s['S'][:N, :N], rng, 0.05) for c in orbs[0]])
    results = {}
    try:
        for mode in (0, 1, 2):
            print(f'case m12u kept-gamma mode {mode} form {form}', flush=True)
            L.uw12g_test_set_kept_gt(mode)
            ctx = u.Context(0)
            W, V, WV, P = gpu_setup(s, ctx)
            results[mode] = [u.form_fock(W, V, WV, P, o, s['occ'], s['n_active'], **kw) for o in orbs]
            del W, V, WV, P
            ctx.close()
    finally:
        L.uw12g_test_set_kept_gt(1)
    for mode in (1, 2):
        for (e, F), (e0, F0) in zip(results[mode], results[0]):
            for k in TERMS + ('total',):
                assert e[k] == e0[k], (mode, k)
            for a, b in zip(F, F0):
                assert np.array_equal(a, b), (mode, np.abs(a - b).max())
