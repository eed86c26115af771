"""c5, the configuration the headline metric is quoted on: 188 atoms, n_ao=1500,
n_df=4500, n_ri=6000, unrestricted 300 / 290 occupied, kappa = (1, 0).  The
integrals are generated on the device and not stored: the Gamma tensors alone
would be 405 GB per operator in FP64, past the 288 GB of one MI355X.

The accuracy gate of the metric: the mixed form and its FP16x3 split variant
against the plain FP64 path of the same run, every energy term within 1e-6 Eh
and the Fock within 1e-6 max-abs."""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, fock_error
from libuw12_amd import synth

pytestmark = pytest.mark.gpu


def test_c5_mixed_gate():
    print('case c5 setup (device-generated, non-stored integrals)', flush=True)
    ctx = u.Context(0)
    s = synth.make_system('c5', with_integrals=False)
    W, V, WV = synth.device_integrals(s, ctx, stored=False)
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=ctx)
    args = (W, V, WV, P, s['orbitals'], s['occ'], s['n_active'])
    kw = dict(indirect_term=False, scale_opp_spin=1.0, scale_same_spin=0.0)
    print('case c5 FP64', flush=True)
    e64, F64 = u.form_fock(*args, **kw, precision=u.FP64)
    for f in F64:
        assert np.array_equal(f, f.T)
    for form, name in ((u.MIXED, 'mixed'), (u.MIXED_FP16X3, 'FP16x3')):
        print(f'case c5 {name}', flush=True)
        em, Fm = u.form_fock(*args, **kw, precision=form)
        for k in TERMS + ('total',):
            print(f'  {k:12s} FP64 {e64[k]: .12f}  {name} - FP64 {em[k] - e64[k]: .2e}', flush=True)
            assert abs(em[k] - e64[k]) <= 1e-6, (name, k, em[k], e64[k])
        d = fock_error(Fm, F64)
        print(f'  Fock max-abs: {name} - FP64 {d:.2e}', flush=True)
        assert 0.0 < d <= 1e-6
    del W, V, WV, P
    ctx.close()
