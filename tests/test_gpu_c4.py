"""c4, the differentiation-form configuration at full size: 100 atoms,
n_ao=800, n_df=2400, n_ri=3200, closed shell with 160 occupied, kappa = (1, 0),
integrals generated on the device and stored.

* the adjoint Fock (uw12g_forward + uw12g_adjoint) equals the fused Fock
  (FP64 to 1e-14 of max|F|; the mixed and FP16x3 adjoints within 1e-6 of the
  FP64 one);
* the adjoint against Richardson-extrapolated central differences of the FP64
  energy along 4 seeded S-consistent orbital paths C(t) = C + t dC,
  dC = (1 - C C^T S) G, G ~ N(0, 1) 1e-4, t = +-1, +-2 (truncation O(t^4)):
  FP64 within 1e-9, the mixed and FP16x3 forms within 1e-6 of the largest |dE/dt|."""
import numpy as np
import pytest

import libuw12_amd as u
from helpers import density_derivative
from libuw12_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def c4():
    print('case c4 setup (device-generated integrals)', flush=True)
    ctx = u.Context(0)
    s = synth.make_system('c4', with_integrals=False)
    W, V, WV = synth.device_integrals(s, ctx, stored=True)
    P = u.setup_abs_projectors(s['S'], s['N'], s['Nr'], context=ctx)
    yield s, (W, V, WV, P)
    del W, V, WV, P
    ctx.close()


def test_c4_adjoint_and_finite_differences(c4):
    s, (W, V, WV, P) = c4
    N = s['N']
    orbs, occ, act = s['orbitals'], s['occ'], s['n_active']
    args = (W, V, WV, P)
    kw = dict(indirect_term=False, scale_opp_spin=1.0, scale_same_spin=0.0)
    F_adj = {}
    for name, prec in (('fp64', u.FP64), ('mixed', u.MIXED), ('fp16x3', u.MIXED_FP16X3)):
        print(f'case c4 {name} forward + adjoint', flush=True)
        _, tape = u.forward(*args, orbs, occ, act, **kw, precision=prec)
        F_adj[name] = u.adjoint(tape)
    _, F_fused = u.form_fock(*args, orbs, occ, act, **kw, precision=u.FP64)
    scale_F = max(np.abs(f).max() for f in F_fused)
    assert max(np.abs(a - b).max() for a, b in zip(F_adj['fp64'], F_fused)) <= 1e-14 * scale_F
    assert max(np.abs(a - b).max() for a, b in zip(F_adj['mixed'], F_fused)) <= 1e-6
    assert max(np.abs(a - b).max() for a, b in zip(F_adj['fp16x3'], F_fused)) <= 1e-6
    Sao = s['S'][:N, :N]
    rng = np.random.default_rng(4004)
    rows = []
    for d in range(4):
        C = orbs[0][:, :len(occ[0])]
        G = rng.standard_normal(C.shape) * 1e-4
        dC = np.zeros_like(orbs[0])
        dC[:, :C.shape[1]] = G - C @ (C.T @ (Sao @ G))
        E = {}
        for t in (1.0, -1.0, 2.0, -2.0):
            e, _ = u.form_fock(*args, [np.asfortranarray(orbs[0] + t * dC)], occ, act, calculate_fock=False, **kw)
            E[t] = e['total']
        fd = (4.0 * 0.5 * (E[1.0] - E[-1.0]) - 0.25 * (E[2.0] - E[-2.0])) / 3.0
        an = {k: density_derivative(F, orbs, occ, [dC]) for k, F in F_adj.items()}
        print(f'  direction {d}: fd {fd:.12e}  fp64 {an["fp64"] - fd: .2e}  mixed {an["mixed"] - fd: .2e}  '
              f'fp16x3 {an["fp16x3"] - fd: .2e}', flush=True)
        rows.append((fd, an))
    scale = max(abs(fd) for fd, _ in rows)
    assert max(abs(an['fp64'] - fd) for fd, an in rows) <= 1e-9 * scale
    assert max(abs(an['mixed'] - fd) for fd, an in rows) <= 1e-6 * scale
    assert max(abs(an['fp16x3'] - fd) for fd, an in rows) <= 1e-6 * scale
