"""Differentiation form: uw12g_forward / uw12g_adjoint through the C ABI, and a
torch.autograd.Function built on them (test harness only).  The adjoint Fock is
checked against the fused Fock, against central finite differences of the
energy along S-consistent density perturbations, and with torch's gradcheck."""
import numpy as np
import pytest
import torch

import libuw12_amd as u
from helpers import gpu_setup
from libuw12_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def sys_open(gpu_ctx):
    s = synth.random_system(9, 12, 14, [4, 3], None, seed=211)
    return s, gpu_setup(s, gpu_ctx)


@pytest.mark.parametrize('prec', [u.FP64, u.FP32, u.MIXED, u.MIXED_FP16X3])
@pytest.mark.parametrize('kappa', [(1.0, 0.5), (1.0, 0.0)])
def test_adjoint_equals_fused(sys_open, prec, kappa):
    s, h = sys_open
    W, V, WV, P = h
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    ef, Ff = u.form_fock(*args, indirect_term=kappa[1] != 0, scale_opp_spin=kappa[0], scale_same_spin=kappa[1],
                         precision=prec)
    e, tape = u.forward(*args, indirect_term=kappa[1] != 0, scale_opp_spin=kappa[0], scale_same_spin=kappa[1],
                        precision=prec)
    F = u.adjoint(tape)
    # the forward forms the same-spin 3e energy from GW/GV, the fused path from its
    # Fock intermediate L: equal to rounding; every other piece bitwise
    for k in u.TERMS:
        if k == 'e3_indirect':  # reduced forms: the fused path reads it off the FP32 L intermediate
            bar = {u.FP64: 1e-12, u.FP32: 1e-5, u.MIXED: 1e-6, u.MIXED_FP16X3: 1e-6}[prec]
            assert abs(e[k] - ef[k]) <= bar * max(1.0, abs(ef[k]))
        else:
            assert e[k] == ef[k], k
    for a, b in zip(F, Ff):
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
    F2 = u.adjoint(tape)  # the adjoint leaves the tape valid
    assert all(np.array_equal(a, b) for a, b in zip(F, F2))


def test_tape_superseded(sys_open):
    s, h = sys_open
    W, V, WV, P = h
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    _, tape = u.forward(*args, scale_same_spin=0.5)
    u.form_fock(*args, scale_same_spin=0.5)  # a later evaluation replaces the intermediates
    with pytest.raises(u.UW12Error) as err:
        u.adjoint(tape)
    assert err.value.code == -4
    _, tape2 = u.forward(*args, scale_same_spin=0.5)
    with pytest.raises(u.UW12Error):
        u.adjoint(tape)  # superseded by tape2
    u.adjoint(tape2)


class UW12Energy(torch.autograd.Function):
    """E(C_alpha, C_beta) with dE/dC_s = 2 F_s C_s[:, :o] diag(occ_s) from the adjoint
    (D_s = C_s,occ diag(occ_s) C_s,occ^T, F_s = dE/dD_s).  When the forward
    intermediates have been replaced (e.g. by gradcheck's finite differences),
    backward recomputes the forward first."""

    @staticmethod
    def forward(ctx, h, s, kappa, *C):
        W, V, WV, P = h
        orbs = [c.detach().cpu().numpy() for c in C]
        e, tape = u.forward(W, V, WV, P, orbs, s['occ'], s['n_active'], indirect_term=kappa[1] != 0,
                            scale_opp_spin=kappa[0], scale_same_spin=kappa[1])
        ctx.h, ctx.s, ctx.kappa, ctx.tape, ctx.orbs = h, s, kappa, tape, orbs
        return torch.tensor(e['total'], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        try:
            F = u.adjoint(ctx.tape)
        except u.UW12Error as err:
            if err.code != -4:
                raise
            W, V, WV, P = ctx.h
            _, tape = u.forward(W, V, WV, P, ctx.orbs, ctx.s['occ'], ctx.s['n_active'],
                                indirect_term=ctx.kappa[1] != 0, scale_opp_spin=ctx.kappa[0],
                                scale_same_spin=ctx.kappa[1])
            F = u.adjoint(tape)
        grads = []
        for Fs, Cs, occ in zip(F, ctx.orbs, ctx.s['occ']):
            G = np.zeros_like(Cs)
            o = len(occ)
            G[:, :o] = 2.0 * Fs @ Cs[:, :o] * occ[None, :]
            grads.append(gE * torch.from_numpy(G))
        return (None, None, None, *grads)


def test_autograd_gradcheck(gpu_ctx):
    s = synth.random_system(6, 9, 10, [3, 2], None, seed=223)
    h = gpu_setup(s, gpu_ctx)
    C = [torch.tensor(c, dtype=torch.float64, requires_grad=True) for c in s['orbitals']]
    fn = lambda *c: UW12Energy.apply(h, s, (1.0, 0.5), *c)  # noqa: E731
    assert torch.autograd.gradcheck(fn, tuple(C), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize('n_occ', [[4], [4, 3]])
def test_adjoint_density_fd(gpu_ctx, n_occ):
    """Central differences of E along D(t) = (C + t dC) occ (C + t dC)^T with the
    S-consistent direction dC = (1 - C C^T S) G against <F, dD/dt>."""
    s = synth.random_system(9, 12, 14, n_occ, None, seed=227)
    h = gpu_setup(s, gpu_ctx)
    W, V, WV, P = h
    S_ao = s['S'][:s['N'], :s['N']]
    rng = np.random.default_rng(5)
    _, tape = u.forward(W, V, WV, P, s['orbitals'], s['occ'], s['n_active'], scale_same_spin=0.5)
    F = u.adjoint(tape)
    for _ in range(3):
        dC = []
        for c in s['orbitals']:
            G = rng.standard_normal(c.shape)
            d = (np.eye(c.shape[0]) - c @ c.T @ S_ao) @ G
            dC.append(d / np.abs(d).max())  # unit size, so that h = 1e-4 stays in the linear range

        def E(t):
            C = [c + t * d for c, d in zip(s['orbitals'], dC)]
            return u.forward(W, V, WV, P, C, s['occ'], s['n_active'], scale_same_spin=0.5)[0]['total']

        hstep = 1e-4
        fd = (E(hstep) - E(-hstep)) / (2 * hstep)
        an = 0.0
        for Fs, c, d, occ in zip(F, s['orbitals'], dC, s['occ']):
            o = len(occ)
            dD = d[:, :o] @ np.diag(occ) @ c[:, :o].T + c[:, :o] @ np.diag(occ) @ d[:, :o].T
            an += np.sum(Fs * dD)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


@pytest.mark.parametrize('kappa', [(1.0, 0.5), (1.0, 0.0)])
def test_recompute_tape_matches_kept_tape(gpu_ctx, kappa):
    """With the tape cap forced below the intermediates' size the forward keeps nothing and
    the adjoint recomputes them: the same energies and the same Fock as the kept tape, and
    the tape stays valid across later evaluations."""
    print(f'case recompute tape m6u kappa={kappa}', flush=True)
    s = synth.make_system('m6u')
    W, V, WV, P = gpu_setup(s, gpu_ctx)
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    kw = dict(indirect_term=kappa[1] != 0, scale_opp_spin=kappa[0], scale_same_spin=kappa[1])
    e_kept, tape = u.forward(*args, **kw)
    F_kept = u.adjoint(tape)
    L = u._lib.load()
    u._lib.check(L.uw12g_test_set_tape_cap(gpu_ctx._h, 1))
    try:
        e_rec, tape_r = u.forward(*args, **kw)
        u.form_fock(*args, **kw)  # an evaluation in between leaves a recompute tape valid
        F_rec = u.adjoint(tape_r)
    finally:
        u._lib.check(L.uw12g_test_set_tape_cap(gpu_ctx._h, 0))
    for k in u.TERMS:
        assert e_rec[k] == e_kept[k], k
    for a, b in zip(F_rec, F_kept):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_streamed_forward_adjoint(gpu_ctx):
    """Non-stored integrals: the forward keeps nothing; the adjoint regenerates the
    three-index blocks (a streamed fused evaluation) and equals the stored adjoint."""
    print('case streamed forward / adjoint m6u', flush=True)
    s = synth.make_system('m6u')
    Ws, Vs, WVs = synth.device_integrals(s, gpu_ctx, stored=False)
    W, V, WV = synth.device_integrals(s, gpu_ctx, stored=True)
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=gpu_ctx)
    orb = (s['orbitals'], s['occ'], s['n_active'])
    kw = dict(indirect_term=False, scale_opp_spin=1.0, scale_same_spin=0.0)
    e_s, tape_s = u.forward(Ws, Vs, WVs, P, *orb, **kw)
    F_s = u.adjoint(tape_s)
    e_f, F_f = u.form_fock(Ws, Vs, WVs, P, *orb, **kw)
    for k in u.TERMS:
        assert e_s[k] == e_f[k], k
    assert all(np.array_equal(a, b) for a, b in zip(F_s, F_f))
    e_0, tape_0 = u.forward(W, V, WV, P, *orb, **kw)
    F_0 = u.adjoint(tape_0)
    for k in u.TERMS:
        assert abs(e_s[k] - e_0[k]) <= 1e-11 * max(1.0, abs(e_0[k])), k
    for a, b in zip(F_s, F_0):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
