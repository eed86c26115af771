"""c3, the precision-form configuration at full size: 100 atoms, n_ao=800,
n_df=2400, n_ri=3200, unrestricted 160 / 150 occupied (beta a seeded rotation
of alpha), kappa = (1, 0), integrals generated on the device and stored.

* mixed form against the plain FP64 path of the same run: every energy term
  within 1e-6 Eh and the Fock within 1e-6 max-abs (the accuracy gate of the
  headline metric);
* FP32 form: each term within 1e-6 of its size (e2 direct is a ~10^2 Eh term)
  and the Fock within 1e-4 -- its errors are reported (measured ~2e-5 Eh on
  e2 / e3 direct, ~1e-5 on the Fock);
* BF16 split forms (the mixed form with its Fock intermediates as BF16x6 /
  BF16x3 products): energies within 1e-6 Eh of FP64 (their paths stay FP64
  as in the mixed form); BF16x6 Fock within the 1e-6 gate,
  BF16x3 Fock within 1e-4, errors reported."""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error
from libuw12_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def c3():
    print('case c3 setup (device-generated integrals)', flush=True)
    ctx = u.Context(0)
    s = synth.make_system('c3', with_integrals=False)
    W, V, WV = synth.device_integrals(s, ctx, stored=True)
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=ctx)
    yield s, (W, V, WV, P)
    del W, V, WV, P
    ctx.close()


def test_c3_precision_forms(c3):
    s, (W, V, WV, P) = c3
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    kw = dict(indirect_term=False, scale_opp_spin=1.0, scale_same_spin=0.0)
    print('case c3 FP64', flush=True)
    e64, F64 = u.form_fock(*args, **kw, precision=u.FP64)
    print('case c3 mixed', flush=True)
    em, Fm = u.form_fock(*args, **kw, precision=u.MIXED)
    print('case c3 FP32', flush=True)
    e32, F32 = u.form_fock(*args, **kw, precision=u.FP32)
    for k in TERMS + ('total',):
        print(f'  {k:12s} FP64 {e64[k]: .12f}  mixed {em[k] - e64[k]: .2e}  FP32 {e32[k] - e64[k]: .2e}', flush=True)
        assert abs(em[k] - e64[k]) <= 1e-6, (k, em[k], e64[k])
        assert abs(e32[k] - e64[k]) <= 1e-6 * max(1.0, abs(e64[k])), (k, e32[k], e64[k])
    dm, d32 = fock_error(Fm, F64), fock_error(F32, F64)
    print(f'  Fock max-abs: mixed {dm:.2e}  FP32 {d32:.2e}  (max|F| {max(np.abs(f).max() for f in F64):.3f})',
          flush=True)
    assert 0.0 < dm <= 1e-6
    assert d32 <= 1e-4
    for form, name, bar in ((u.MIXED_BF16X6, 'BF16x6', 1e-6), (u.MIXED_BF16X3, 'BF16x3', 1e-4),
                            (u.MIXED_FP16X3, 'FP16x3', 1e-6)):
        print(f'case c3 {name}', flush=True)
        eb, Fb = u.form_fock(*args, **kw, precision=form)
        for k in TERMS + ('total',):
            assert abs(eb[k] - e64[k]) <= 1e-6, (name, k, eb[k], e64[k])
        db = fock_error(Fb, F64)
        print(f'  Fock max-abs: {name} {db:.2e}', flush=True)
        assert 0.0 < db <= bar
