"""GPU path: the reference test-suite relations, error handling, finite
differences, determinism, and parity against the oracle on the s-Gaussian
recipe at intermediate sizes and on the reference's water integrals."""
import os

import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error, gpu_eval, gpu_setup, oracle_eval
from libuw12_amd import synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def rsys(gpu_ctx):
    s = synth.random_system(9, 12, 14, [4, 3], [3, 2], seed=101)
    return s, gpu_setup(s, gpu_ctx)


def test_linear_in_scale_factors(rsys):
    s, h = rsys
    e, F = gpu_eval(h, s, 1.0, 0.5)
    eos, Fos = gpu_eval(h, s, 1.0, 0.0, indirect=True)
    ess, Fss = gpu_eval(h, s, 0.0, 1.0)
    assert abs(eos['total'] + 0.5 * ess['total'] - e['total']) <= 1e-12 * abs(e['total'])
    for a, b, c in zip(Fos, Fss, F):
        assert np.abs(a + 0.5 * b - c).max() <= 1e-12 * np.abs(c).max()


def test_w_v_symmetry(gpu_ctx, rsys):
    s, h = rsys
    W, V, WV, P = h
    e1, _ = gpu_eval(h, s, 1.0, 0.5)
    e2, _ = gpu_eval((V, W, WV, P), s, 1.0, 0.5)
    for k in ('e3_direct', 'e3_indirect', 'e4_direct', 'e4_indirect'):
        assert abs(e1[k] - e2[k]) <= 1e-12 * max(abs(e1[k]), 1.0), k


def test_open_shell_equal_orbitals_is_closed_shell(gpu_ctx):
    s = synth.random_system(9, 12, 14, [4], None, seed=103)
    h = gpu_setup(s, gpu_ctx)
    C, occ = s['orbitals'][0], s['occ'][0]
    ec, Fc = gpu_eval(h, s, 1.0, 0.5)
    eo, Fo = gpu_eval(h, s, 1.0, 0.5, orbitals=[C, C], occ=[occ / 2, occ / 2], n_active=[4, 4])
    assert abs(ec['total'] - eo['total']) <= 1e-12 * abs(ec['total'])
    for f in Fo:
        assert np.abs(f - Fc[0]).max() <= 1e-11 * np.abs(Fc[0]).max()


def test_no_active_orbitals_gives_zero(rsys):
    s, h = rsys
    e, F = gpu_eval(h, s, 1.0, 0.5, n_active=[0, 0])
    assert all(e[k] == 0.0 for k in TERMS)
    assert all(np.abs(f).max() == 0.0 for f in F)


def test_energy_only_and_indirect_flag(rsys):
    s, h = rsys
    e1, F1 = gpu_eval(h, s, 1.0, 0.5)
    e2, F2 = gpu_eval(h, s, 1.0, 0.5, fock=False)
    assert F2 is None
    # with the Fock the same-spin 3e energy is read off the Fock's L intermediate,
    # without it from the GW/GV blocks: equal up to rounding, every other piece bitwise
    for k in TERMS:
        if k == 'e3_indirect':
            assert abs(e1[k] - e2[k]) <= 1e-12 * abs(e2[k])
        else:
            assert e1[k] == e2[k], k
    assert abs(e1['total'] - e2['total']) <= 1e-12 * abs(e2['total'])
    a, Fa = gpu_eval(h, s, 1.0, 0.0, indirect=True)
    b, Fb = gpu_eval(h, s, 1.0, 0.0, indirect=False)
    assert a['total'] == b['total']
    assert all(np.array_equal(x, y) for x, y in zip(Fa, Fb))


def test_deterministic(rsys):
    s, h = rsys
    e1, F1 = gpu_eval(h, s, 1.0, 0.5)
    e2, F2 = gpu_eval(h, s, 1.0, 0.5)
    assert all(e1[k] == e2[k] for k in TERMS)
    assert all(np.array_equal(x, y) for x, y in zip(F1, F2))


def test_one_electron_self_interaction(gpu_ctx):
    s = synth.random_system(9, 12, 14, [1, 0], None, seed=107)
    h = gpu_setup(s, gpu_ctx)
    e, F = gpu_eval(h, s, 1.0, 1.0, occ=[np.ones(1), np.ones(0)], n_active=[1, 0])
    assert abs(e['e2_direct'] + e['e2_indirect']) <= 1e-12 * abs(e['e2_direct'])
    eo, Fo = oracle_eval(s, 1.0, 1.0, occ=[np.ones(1), np.ones(0)], n_active=[1, 0])
    for k in TERMS:
        assert abs(e[k] - eo[k]) <= 1e-12 * max(abs(eo[k]), 1.0), k
    assert fock_error(F, Fo) <= 1e-10 * max(np.abs(f).max() for f in Fo)


@pytest.mark.parametrize('n_occ', [[4], [4, 3]])
def test_fock_is_energy_derivative(gpu_ctx, n_occ):
    s = synth.random_system(9, 12, 14, n_occ, None, seed=109)
    h = gpu_setup(s, gpu_ctx)
    n_spin = len(n_occ)
    g = 2.0 if n_spin == 1 else 1.0
    _, F = gpu_eval(h, s, 1.0, 0.5)
    rng = np.random.default_rng(3)
    for _ in range(2):
        dC = [rng.standard_normal(c.shape) for c in s['orbitals']]

        def E(t):
            C = [c + t * d for c, d in zip(s['orbitals'], dC)]
            return gpu_eval(h, s, 1.0, 0.5, orbitals=C, fock=False)[0]['total']

        hstep = 1e-4
        fd = (E(hstep) - E(-hstep)) / (2 * hstep)
        an = 0.0
        for sp in range(n_spin):
            o = len(s['occ'][sp])
            w = np.sqrt(s['occ'][sp] / g)
            Co, dCo = s['orbitals'][sp][:, :o] * w, dC[sp][:, :o] * w
            an += np.sum(F[sp] * g * (dCo @ Co.T + Co @ dCo.T))
        assert abs(fd - an) <= 1e-6 * abs(an)


def test_error_handling(gpu_ctx, rsys):
    s, h = rsys
    W, V, WV, P = h
    with pytest.raises(u.UW12Error):  # n_active > n_occ
        gpu_eval(h, s, 1.0, 0.5, n_active=[5, 2])
    with pytest.raises(u.UW12Error):  # non-positive occupation
        gpu_eval(h, s, 1.0, 0.5, occ=[np.array([1.0, 0.0, 1.0, 1.0]), s['occ'][1]])
    with pytest.raises(u.UW12Error):  # spin channel mismatch
        u.form_fock(W, V, WV, P, s['orbitals'], s['occ'][:1], s['n_active'])
    with pytest.raises(u.UW12Error):  # W without RI integrals
        u.form_fock(WV, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    X3, X2, X3ri = s['ints']['W']
    bad = np.array(X2, copy=True)
    bad[0, 1] += 1e-6
    with pytest.raises(u.UW12Error):  # metric not symmetric
        u.setup_base_integrals(X3, bad, X3ri, n_ao=s['N'], context=gpu_ctx)
    with pytest.raises(u.UW12Error):  # n_ao mismatch between integrals and projectors
        small = synth.random_system(8, 12, 14, [4, 3], None, seed=5)
        P2 = u.setup_abs_projectors(small['S'], 8, 14, context=gpu_ctx)
        u.form_fock(W, V, WV, P2, s['orbitals'], s['occ'], s['n_active'])
    # the context stays usable after errors
    e, _ = gpu_eval(h, s, 1.0, 0.5)
    assert np.isfinite(e['total'])


def test_abs_projector_matches_oracle(gpu_ctx):
    from oracle import cpu
    s = synth.make_system('c1')
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=gpu_ctx)
    Sg = P.matrix()
    So = cpu.abs_projector(s['S'], s['N'], s['Nr'])
    # S+ has norm ~2e5 here (kept eigenvalues down to ~1e-6): two FP64 eigensolvers
    # differ by ~|S+|^2 eps |S| in it, ~2e-9 relative on this recipe
    assert np.abs(Sg - So).max() <= 1e-8 * np.abs(So).max()


@pytest.mark.parametrize('name', ['m6', 'm6u'])
@pytest.mark.parametrize('kappa', [(1.0, 0.5), (1.0, 0.0)])
def test_recipe_mid_size(gpu_ctx, name, kappa):
    print(f'case {name} kappa={kappa}', flush=True)
    s = synth.make_system(name)
    h = gpu_setup(s, gpu_ctx)
    eg, Fg = gpu_eval(h, s, *kappa)
    eo, Fo = oracle_eval(s, *kappa)
    for k in TERMS + ('total',):
        assert abs(eg[k] - eo[k]) <= 2e-9 * max(1.0, abs(eo[k])), (k, eg[k], eo[k])
    assert fock_error(Fg, Fo) <= 2e-8 * max(1.0, max(np.abs(f).max() for f in Fo))


@pytest.mark.parametrize('name,n_active', [('m12', None), ('m12', [19]), ('m12u', None)])
def test_recipe_m12_same_spin(gpu_ctx, name, n_active):
    """24 occupied: the second same-spin pass pairs its active j in blocks of 16, with the sums S
    of the pairs past the first block; with 19 active, the 5 core k join every block; open shell
    (24 / 23) pairs each spin on its own."""
    print(f'case {name} kappa=(1.0, 0.5) n_active={n_active}', flush=True)
    s = synth.make_system(name)
    h = gpu_setup(s, gpu_ctx)
    eg, Fg = gpu_eval(h, s, 1.0, 0.5, n_active=n_active)
    eo, Fo = oracle_eval(s, 1.0, 0.5, n_active=n_active)
    for k in TERMS + ('total',):
        assert abs(eg[k] - eo[k]) <= 2e-9 * max(1.0, abs(eo[k])), (k, eg[k], eo[k])
    assert fock_error(Fg, Fo) <= 2e-8 * max(1.0, max(np.abs(f).max() for f in Fo))


def test_recipe_c2_opposite_spin(gpu_ctx):
    """Full c2 shape (n_ao=320, n_df=960, n_ri=1280, n_occ=80) at kappa = (1, 0) against the live
    oracle, with absolute bars: 2e / 4e terms 1e-10 Eh, the 3e terms, total and Fock 3e-8 (the
    oracle's explicit-S+ rounding on this recipe; tests/test_gpu_c2_parity.py)."""
    print('case c2 kappa=(1.0, 0.0)', flush=True)
    s = synth.make_system('c2')
    h = gpu_setup(s, gpu_ctx)
    eg, Fg = gpu_eval(h, s, 1.0, 0.0)
    eo, Fo = oracle_eval(s, 1.0, 0.0)
    for k in TERMS + ('total',):
        bar = 3e-8 if k in ('e3_direct', 'e3_indirect', 'total') else 1e-10
        assert abs(eg[k] - eo[k]) <= bar, (k, eg[k], eo[k])
    assert fock_error(Fg, Fo) <= 3e-8


def test_water_integrals(gpu_ctx):
    """The reference's water DF integrals (n_ao=24, n_df=116, n_ri=76) with a
    zero RI block (W3_ri/V3_ri are not in the reference checkout), frozen core."""
    w = dict(np.load(os.path.join(HERE, 'golden', 'water.npz')))
    N, Nd, Nr = 24, w['W2'].shape[0], w['S'].shape[0] - 24
    ri = np.zeros((N * Nr, Nd), order='F')
    for spin in ('neutral', 'cation'):
        if spin == 'neutral':
            orbs, occ, act = [w['orbitals']], [w['occ'][:, 0]], [4]
        else:
            orbs, occ, act = [w['orbitals_alpha'], w['orbitals_beta']], [w['occ_alpha'][:, 0], w['occ_beta'][:, 0]], [4, 3]
        s = dict(N=N, Nd=Nd, Nr=Nr, S=w['S'], orbitals=orbs, occ=occ, n_active=act,
                 ints={'W': (w['W3'], w['W2'], ri), 'V': (w['V3'], w['V2'], ri), 'WV': (w['WV3'], w['WV2'], None)})
        h = gpu_setup(s, gpu_ctx)
        eg, Fg = gpu_eval(h, s, 1.0, 0.5)
        eo, Fo = oracle_eval(s, 1.0, 0.5)
        # with the RI block zeroed the three-electron term is unphysically large
        # (~2e5 Eh); the bar is relative to each term's size
        for k in TERMS + ('total',):
            assert abs(eg[k] - eo[k]) <= 1e-9 * max(1.0, abs(eo[k])), (spin, k, eg[k], eo[k])
        assert fock_error(Fg, Fo) <= 1e-8 * max(1.0, max(np.abs(f).max() for f in Fo))


@pytest.mark.parametrize('name', ['c1', 'm6u', 'm12'])
@pytest.mark.parametrize('kappa', [(1.0, 0.5), (1.0, 0.0)])
def test_mixed_form_errors(gpu_ctx, name, kappa):
    """Mixed form against FP64: the energy paths (orbital transforms, t, X,
    energy-only products) and every accumulation into F stay FP64, so all energy
    terms but the same-spin 3e one are bit-identical; that one is read off the
    FP32 L intermediate.  Bar: 1e-6 Eh per term and 1e-6 Fock max-abs (measured
    <= 3e-7 and <= 4e-7 on these systems)."""
    print(f'case {name} kappa={kappa} mixed form', flush=True)
    s = synth.make_system(name)
    W, V, WV, P = gpu_setup(s, gpu_ctx)
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    kw = dict(indirect_term=True, scale_opp_spin=kappa[0], scale_same_spin=kappa[1])
    e64, F64 = u.form_fock(*args, **kw, precision=u.FP64)
    em, Fm = u.form_fock(*args, **kw, precision=u.MIXED)
    for k in TERMS:
        if k == 'e3_indirect' and kappa[1] != 0.0:
            assert abs(em[k] - e64[k]) <= 1e-6, (k, em[k], e64[k])
        else:
            assert em[k] == e64[k], (k, em[k], e64[k])
    assert abs(em['total'] - e64['total']) <= 1e-6
    assert 0.0 < fock_error(Fm, F64) <= 1e-6


@pytest.mark.parametrize('name', ['c1', 'm6u', 'm12'])
@pytest.mark.parametrize('kappa', [(1.0, 0.5), (1.0, 0.0)])
def test_bf16_split_forms_errors(gpu_ctx, name, kappa):
    """Mixed form with the Fock intermediates as BF16 split products, against FP64.
    As in the mixed form, every energy term but the same-spin 3e one is
    bit-identical.  Three planes (per-product error ~2^-22, below FP32's
    rounding of the operands): the accuracy gate, 1e-6 Eh per term and 1e-6
    Fock max-abs.  Two planes (~3 * 2^-16 per product): e3 indirect within 1e-5
    Eh and the Fock within 5e-5 (measured 7e-7 and 5e-6 at n_ao = 320), and no
    more accurate than the three-plane form.  Two half planes of row-scaled
    operands (FP16x3, ~3 * 2^-22 per product): the accuracy gate, and more
    accurate than the two-plane BF16 form."""
    print(f'case {name} kappa={kappa} BF16 / FP16 split forms', flush=True)
    s = synth.make_system(name)
    W, V, WV, P = gpu_setup(s, gpu_ctx)
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    kw = dict(indirect_term=True, scale_opp_spin=kappa[0], scale_same_spin=kappa[1])
    e64, F64 = u.form_fock(*args, **kw, precision=u.FP64)
    errs = {}
    for form, e_bar, f_bar in ((u.MIXED_BF16X6, 1e-6, 1e-6), (u.MIXED_BF16X3, 1e-5, 5e-5),
                               (u.MIXED_FP16X3, 1e-6, 1e-6)):
        e, F = u.form_fock(*args, **kw, precision=form)
        for k in TERMS:
            if k == 'e3_indirect' and kappa[1] != 0.0:
                assert abs(e[k] - e64[k]) <= e_bar, (form, k, e[k], e64[k])
            else:
                assert e[k] == e64[k], (form, k, e[k], e64[k])
        errs[form] = fock_error(F, F64)
        print(f'  form {form}: e3 indirect {e["e3_indirect"] - e64["e3_indirect"]:.2e}  Fock {errs[form]:.2e}',
              flush=True)
        assert 0.0 < errs[form] <= f_bar
    assert errs[u.MIXED_BF16X6] <= errs[u.MIXED_BF16X3]
    assert errs[u.MIXED_FP16X3] <= errs[u.MIXED_BF16X3]


def test_fp32_form_errors(gpu_ctx):
    """FP32 form against FP64 on the c1 recipe: every contraction rounds its
    operands to FP32 and sums in FP32 over k chunks of at most 4096, so each
    energy term moves by ~1e-7 of its size (measured <= 1.1e-7 on c1..m12) and
    the Fock by ~1e-5 absolute (long cancelling sums into F)."""
    print('case c1 fp32 form', flush=True)
    s = synth.make_system('c1')
    h = gpu_setup(s, gpu_ctx)
    W, V, WV, P = h
    e64, F64 = u.form_fock(W, V, WV, P, s['orbitals'], s['occ'], s['n_active'], indirect_term=True,
                           scale_opp_spin=1.0, scale_same_spin=0.5, precision=u.FP64)
    e32, F32 = u.form_fock(W, V, WV, P, s['orbitals'], s['occ'], s['n_active'], indirect_term=True,
                           scale_opp_spin=1.0, scale_same_spin=0.5, precision=u.FP32)
    for k in TERMS + ('total',):
        assert abs(e32[k] - e64[k]) <= 1e-6 * max(1.0, abs(e64[k])), (k, e32[k], e64[k])
        assert e32[k] != e64[k]
    assert fock_error(F32, F64) <= 1e-4
    with pytest.raises(u.UW12Error):
        u.form_fock(W, V, WV, P, s['orbitals'], s['occ'], s['n_active'], precision=7)
