"""Device-generated integrals (uw12g_setup_base_integrals_gen) and the evaluation
over non-stored integrals, which requests every three-index block again from
the device generator and never holds a tensor of the size of X3_ri.

* The device generator reproduces the host generator's blocks (the analytic
  s-Gaussian formulas evaluated with the device math library) to 1e-13
  relative; its AO block is exactly symmetric.
* Stored device-generated integrals give the host-array setup's results to
  FP64 rounding (1e-12 relative per term, 1e-11 of max|F|).
* Non-stored integrals (blocked over the ABS eigenvectors and DF functions)
  give the stored path's opposite-spin terms and Fock to 1e-11 relative /
  1e-10 of max|F| -- the algebra is the same, only the summation order moves
  -- also with blocks forced down to 16 ABS rows and one DF function.
"""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error, fock_scale, gpu_eval, gpu_setup
from libuw12_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('name', ['c1', 'm6'])
@pytest.mark.parametrize('kind', ['W', 'V', 'WV'])
def test_device_generator_matches_host(name, kind):
    print(f'case generator {name} {kind}', flush=True)
    s = synth.make_system(name)
    N, Nr, Nd = s['N'], s['Nr'], s['Nd']
    X3, _, X3ri = s['ints'][kind]
    gen = synth.DeviceGenerator(s, kind)
    with_ri = kind != 'WV'
    rows = N + (Nr if with_ri else 0)
    df0, n = Nd // 3, Nd - Nd // 3
    blk = gen.block(df0, n, with_ri)
    assert blk.shape == (rows, N, n)
    iu, ju = np.tril_indices(N)  # packed row mu (mu + 1) / 2 + nu, nu <= mu
    ref_ao = np.zeros((N, N, n))
    ref_ao[iu, ju, :] = X3[:, df0:df0 + n]
    ref_ao[ju, iu, :] = X3[:, df0:df0 + n]
    scale = np.abs(ref_ao).max()
    assert np.abs(blk[:N] - ref_ao).max() <= 1e-13 * scale
    symmetric = bool(np.array_equal(blk[:N], blk[:N].transpose(1, 0, 2)))
    assert symmetric  # bitwise symmetric AO block
    if with_ri:
        ref_ri = X3ri[:, df0:df0 + n].reshape(Nr, N, n, order='F')
        assert np.abs(blk[N:] - ref_ri).max() <= 1e-13 * np.abs(ref_ri).max()
    gen.close()


@pytest.mark.parametrize('name', ['c1', 'm6u'])
def test_stored_generated_matches_host_setup(gpu_ctx, name):
    print(f'case stored generated {name}', flush=True)
    s = synth.make_system(name)
    h_host = gpu_setup(s, gpu_ctx)
    W, V, WV = synth.device_integrals(s, gpu_ctx, stored=True)
    P = h_host[3]
    eh, Fh = gpu_eval(h_host, s, 1.0, 0.5)
    eg, Fg = gpu_eval((W, V, WV, P), s, 1.0, 0.5)
    for k in TERMS + ('total',):
        assert abs(eg[k] - eh[k]) <= 1e-12 * max(1.0, abs(eh[k])), (k, eg[k], eh[k])
    assert fock_error(Fg, Fh) <= 1e-11 * fock_scale(Fh)


def streamed_vs_stored(ctx, s, kappa, cap=None, precision=u.FP64, n_active=None, fock=True):
    W, V, WV = synth.device_integrals(s, ctx, stored=True)
    Ws, Vs, WVs = synth.device_integrals(s, ctx, stored=False)
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=ctx)
    kw = dict(indirect_term=False, calculate_fock=fock, scale_opp_spin=kappa[0], scale_same_spin=kappa[1])
    act = n_active or s['n_active']
    e0, F0 = u.form_fock(W, V, WV, P, s['orbitals'], s['occ'], act, **kw)
    if cap is not None:
        u._lib.check(u._lib.load().uw12g_test_set_block_cap(ctx._h, cap))
    try:
        e1, F1 = u.form_fock(Ws, Vs, WVs, P, s['orbitals'], s['occ'], act, precision=precision, **kw)
    finally:
        u._lib.check(u._lib.load().uw12g_test_set_block_cap(ctx._h, 0))
    return e0, F0, e1, F1


@pytest.mark.parametrize('name', ['c1', 'm6u', 'm12u'])
@pytest.mark.parametrize('kappa', [(1.0, 0.0), (1.0, 0.5)])
def test_streamed_matches_stored(gpu_ctx, name, kappa):
    """Opposite-spin terms (indirect_term off) of the non-stored path against the stored one."""
    print(f'case streamed {name} kappa={kappa}', flush=True)
    s = synth.make_system(name)
    e0, F0, e1, F1 = streamed_vs_stored(gpu_ctx, s, kappa)
    for k in TERMS + ('total',):
        assert abs(e1[k] - e0[k]) <= 1e-11 * max(1.0, abs(e0[k])), (k, e1[k], e0[k])
    assert fock_error(F1, F0) <= 1e-10 * fock_scale(F0)


@pytest.mark.parametrize('name', ['c1', 'm6u'])
def test_streamed_small_blocks(gpu_ctx, name):
    """Blocks forced down (block cap of 1024 doubles: 16 ABS rows, one DF function per
    generator call) give the same result."""
    print(f'case streamed small blocks {name}', flush=True)
    s = synth.make_system(name)
    e0, F0, e1, F1 = streamed_vs_stored(gpu_ctx, s, (1.0, 0.0), cap=1024)
    for k in TERMS + ('total',):
        assert abs(e1[k] - e0[k]) <= 1e-11 * max(1.0, abs(e0[k])), (k, e1[k], e0[k])
    assert fock_error(F1, F0) <= 1e-10 * fock_scale(F0)


def test_streamed_frozen_core_and_energy_only(gpu_ctx):
    print('case streamed frozen core, energy only', flush=True)
    s = synth.make_system('m6')
    e0, F0, e1, F1 = streamed_vs_stored(gpu_ctx, s, (1.0, 0.0), n_active=[9])
    for k in TERMS + ('total',):
        assert abs(e1[k] - e0[k]) <= 1e-11 * max(1.0, abs(e0[k])), (k, e1[k], e0[k])
    assert fock_error(F1, F0) <= 1e-10 * fock_scale(F0)
    e0, _, e1, F1 = streamed_vs_stored(gpu_ctx, s, (1.0, 0.0), fock=False)
    assert F1 is None
    for k in TERMS + ('total',):
        assert abs(e1[k] - e0[k]) <= 1e-11 * max(1.0, abs(e0[k])), (k, e1[k], e0[k])


def test_streamed_mixed_form(gpu_ctx):
    """Mixed form over non-stored integrals: the t~ dX intermediate Xi on FP32 matrix
    cores, energies and every accumulation into F in FP64: energies bit-identical to the
    FP64 streamed path, Fock within 1e-6."""
    print('case streamed mixed m6u', flush=True)
    s = synth.make_system('m6u')
    _, _, e64, F64 = streamed_vs_stored(gpu_ctx, s, (1.0, 0.0))
    _, _, em, Fm = streamed_vs_stored(gpu_ctx, s, (1.0, 0.0), precision=u.MIXED)
    for k in TERMS:
        assert em[k] == e64[k], (k, em[k], e64[k])
    assert 0.0 < fock_error(Fm, F64) <= 1e-6


def test_streamed_rejects_unsupported(gpu_ctx):
    s = synth.make_system('c1')
    Ws, Vs, WVs = synth.device_integrals(s, gpu_ctx, stored=False)
    W, V, WV = synth.device_integrals(s, gpu_ctx, stored=True)
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=gpu_ctx)
    args = (s['orbitals'], s['occ'], s['n_active'])
    with pytest.raises(u.UW12Error) as ei:  # same-spin terms need stored integrals
        u.form_fock(Ws, Vs, WVs, P, *args, indirect_term=True, scale_opp_spin=1.0, scale_same_spin=0.5)
    assert ei.value.code == -2
    with pytest.raises(u.UW12Error) as ei:
        u.form_fock(Ws, Vs, WVs, P, *args, indirect_term=False, precision=u.FP32)
    assert ei.value.code == -2
    with pytest.raises(u.UW12Error) as ei:  # the differentiation form shares the restriction
        u.forward(Ws, Vs, WVs, P, *args, indirect_term=True, scale_opp_spin=1.0, scale_same_spin=0.5)
    assert ei.value.code == -2
    with pytest.raises(u.UW12Error) as ei:  # stored W / V with a non-stored WV
        u.form_fock(W, V, WVs, P, *args, indirect_term=False)
    assert ei.value.code == -2
    with pytest.raises(u.UW12Error) as ei:  # W and V must agree
        u.form_fock(W, Vs, WV, P, *args, indirect_term=False)
    assert ei.value.code == -2
    # the energy-only entry point (uw12_energy) sets the indirect flag from kappa: kappa1 = 0 runs
    e = u.uw12_energy(Ws, Vs, WVs, P, *args, 1.0, 0.0)
    e_ref = u.uw12_energy(W, V, WV, P, *args, 1.0, 0.0)
    assert abs(e - e_ref) <= 1e-11 * abs(e_ref)
