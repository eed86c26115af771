"""Function-based (per-DF-shell, streamed) integral setup against the array form."""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error
from libuw12_amd import _lib, synth

pytestmark = pytest.mark.gpu


def setup_fn(s, ctx, cap=None):
    if cap is not None:
        _lib.check(_lib.load().uw12g_test_set_stream_cap(ctx._h, cap))
    out = []
    for kind in ('W', 'V', 'WV'):
        X2, sizes, fn, fn_ri = synth.shell_integrals(s, kind)
        out.append(u.setup_base_integrals_fn(X2, fn, fn_ri, sizes, s['N'], s['Nr'] if fn_ri else 0, context=ctx))
    out.append(u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=ctx))
    return out


def setup_arrays(s, ctx):
    N = s['N']
    W = u.setup_base_integrals(*s['ints']['W'], n_ao=N, context=ctx)
    V = u.setup_base_integrals(*s['ints']['V'], n_ao=N, context=ctx)
    WV = u.setup_base_integrals(s['ints']['WV'][0], s['ints']['WV'][1], None, n_ao=N, context=ctx)
    P = u.setup_abs_projectors(s['S'], N, s['Nr'], context=ctx)
    return W, V, WV, P


@pytest.mark.parametrize('name,cap', [('c1', None), ('c1', 1), ('m6u', 20000)])
def test_shell_streamed_setup_matches_arrays(gpu_ctx, name, cap):
    """cap = 1 streams one shell per group; 20000 doubles groups a few shells."""
    print(f'case {name} cap={cap}', flush=True)
    s = synth.make_system(name)
    ha = setup_arrays(s, gpu_ctx)
    hf = setup_fn(s, gpu_ctx, cap)
    try:
        ea, Fa = u.form_fock(*ha, s['orbitals'], s['occ'], s['n_active'], scale_same_spin=0.5)
        ef, Ff = u.form_fock(*hf, s['orbitals'], s['occ'], s['n_active'], scale_same_spin=0.5)
    finally:
        _lib.check(_lib.load().uw12g_test_set_stream_cap(gpu_ctx._h, 1 << 25))
    # Gamma = G0 T accumulates over DF column groups in a different order: FP64 rounding only
    for k in TERMS + ('total',):
        assert abs(ef[k] - ea[k]) <= 1e-10 * max(1.0, abs(ea[k])), (k, ef[k], ea[k])
    assert fock_error(Ff, Fa) <= 1e-9 * max(np.abs(f).max() for f in Fa)


def test_callback_errors(gpu_ctx):
    s = synth.make_system('c1')
    X2, sizes, fn, fn_ri = synth.shell_integrals(s, 'W')

    def bad(shell):
        if shell == 1:
            raise RuntimeError('integral engine failed')
        return fn(shell)

    with pytest.raises(u.UW12Error) as err:
        u.setup_base_integrals_fn(X2, bad, fn_ri, sizes, s['N'], s['Nr'], context=gpu_ctx)
    assert err.value.code == -6 and 'integral engine failed' in str(err.value)
    with pytest.raises(u.UW12Error):  # wrong block shape from the callback
        u.setup_base_integrals_fn(X2, lambda sh: fn(sh)[:, :1], fn_ri, sizes, s['N'], s['Nr'], context=gpu_ctx)
    with pytest.raises(u.UW12Error):  # shell sizes do not add up to n_df
        u.setup_base_integrals_fn(X2, fn, fn_ri, sizes[:-1], s['N'], s['Nr'], context=gpu_ctx)
    # the context is still usable
    h = setup_fn(s, gpu_ctx)
    e, _ = u.form_fock(*h, s['orbitals'], s['occ'], s['n_active'], scale_same_spin=0.5)
    assert np.isfinite(e['total'])
