"""Blocked same-spin paths against the unblocked ones.

A small block cap (uw12g_test_set_block_cap) splits K~ into one j per block,
the L contraction into column blocks of the DF index (each re-forming K~), the
4e X_i / Y_i intermediates into i (and, with frozen core, l) blocks, Xi into
narrow A blocks and the energy-only GW / GV products into k blocks.  The sums
are the same in real arithmetic: FP64 results agree to rounding."""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error, gpu_setup
from libuw12_amd import _lib, synth

pytestmark = pytest.mark.gpu


def set_cap(ctx, doubles):
    _lib.check(_lib.load().uw12g_test_set_block_cap(ctx._h, doubles))


@pytest.mark.parametrize('name,n_active,cap', [('m6u', None, 20000), ('c1', [4], 2000), ('c1', None, 600)])
@pytest.mark.parametrize('precision', [u.FP64, u.MIXED, u.MIXED_FP16X3])
def test_blocked_equals_unblocked(gpu_ctx, name, n_active, cap, precision):
    print(f'case {name} n_active={n_active} cap={cap} precision={precision}', flush=True)
    s = synth.make_system(name)
    h = gpu_setup(s, gpu_ctx)
    act = n_active or s['n_active']
    kw = dict(indirect_term=True, scale_opp_spin=1.0, scale_same_spin=0.5, precision=precision)
    e0, F0 = u.form_fock(*h, s['orbitals'], s['occ'], act, **kw)
    q0, _ = u.form_fock(*h, s['orbitals'], s['occ'], act, calculate_fock=False, **kw)
    set_cap(gpu_ctx, cap)
    try:
        e1, F1 = u.form_fock(*h, s['orbitals'], s['occ'], act, **kw)
        q1, _ = u.form_fock(*h, s['orbitals'], s['occ'], act, calculate_fock=False, **kw)
    finally:
        set_cap(gpu_ctx, 0)
    # FP64: re-association only; mixed forms: the FP32 k ranges of L follow the j-blocks (the FP16x3
    # form also keeps the blocked K~ in FP32 storage and takes its row scales from bounds)
    rel = 1e-11 if precision == u.FP64 else 1e-6
    scale = max(abs(e0[k]) for k in TERMS)
    for k in TERMS + ('total',):
        assert abs(e1[k] - e0[k]) <= rel * max(abs(e0[k]), 1e-3 * scale), (k, e1[k], e0[k])
        assert abs(q1[k] - q0[k]) <= rel * max(abs(q0[k]), 1e-3 * scale), (k, q1[k], q0[k])
    fs = max(float(np.abs(f).max()) for f in F0)
    assert fock_error(F1, F0) <= (1e-10 if precision == u.FP64 else 2e-6) * fs


def test_progress_lines(gpu_ctx, capfd):
    """uw12g_context_set_verbosity: one line per major stage on stderr, the results unchanged."""
    print('case c1 progress lines', flush=True)
    s = synth.make_system('c1')
    h = gpu_setup(s, gpu_ctx)
    e0, F0 = u.form_fock(*h, s['orbitals'], s['occ'], s['n_active'], scale_same_spin=0.5)
    capfd.readouterr()
    gpu_ctx.set_verbosity(1)
    try:
        e1, F1 = u.form_fock(*h, s['orbitals'], s['occ'], s['n_active'], scale_same_spin=0.5)
    finally:
        gpu_ctx.set_verbosity(0)
    err = capfd.readouterr().err
    for stage in ('orbital transforms', 'direct Fock', '4e indirect Fock', 'same-spin 3e Fock'):
        assert f'{stage}' in err, err
    assert all(line.startswith('[uw12g]') for line in err.splitlines() if 'uw12g' in line)
    assert all(e0[k] == e1[k] for k in TERMS)
    assert all(np.array_equal(a, b) for a, b in zip(F0, F1))
