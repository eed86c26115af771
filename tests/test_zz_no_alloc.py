"""uw12g_form_fock_noalloc (no device allocation: every buffer carved from the
library's static scratch array) against uw12g_form_fock, through the C ABI.

Two calls in a row, the second on a smaller system, each compared with the
allocating entry point on the same inputs: the same kernels in the same order,
so the energies and Fock matrices agree to the last bit.  The first call runs
on fresh integral handles, so U^T Gamma is formed in the scratch array; the
allocating call afterwards keeps it on the handles.  A system whose working
set exceeds the array is refused with -2 and a message naming the size."""
import ctypes

import numpy as np
import pytest

import libuw12_amd as u
from helpers import TERMS, gpu_setup
from libuw12_amd import _lib, synth
from libuw12_amd.api import _orbital_args, _ptr

pytestmark = pytest.mark.gpu


def call(entry, handles, system, kappa, ctx):
    W, V, WV, P = handles
    C, n_ao, n_orb, n_spin, occv, n_occ, n_act = _orbital_args(system['orbitals'], system['occ'],
                                                               system['n_active'])
    terms = np.zeros(6)
    energy = ctypes.c_double(0.0)
    F = np.zeros((n_ao, n_ao * n_spin), order='F')
    rc = getattr(_lib.load(), entry)(ctx._h, W._h, V._h, WV._h, P._h, _ptr(C), n_ao, n_orb, n_spin, _ptr(occv), n_occ,
                                     n_act, 1, float(kappa[0]), float(kappa[1]), u.FP64, _ptr(F), _ptr(terms),
                                     ctypes.byref(energy))
    return rc, terms, energy.value, F


def test_noalloc_matches_form_fock_twice():
    ctx = u.Context(0)
    for name, kappa in (('m6u', (1.0, 0.5)), ('c1', (1.0, 0.5))):  # the second input is smaller
        print(f'case noalloc {name} kappa={kappa}', flush=True)
        s = synth.make_system(name)
        h = gpu_setup(s, ctx)
        rc, t1, e1, F1 = call('uw12g_form_fock_noalloc', h, s, kappa, ctx)
        assert rc == 0, _lib.load().uw12g_last_error().decode()
        rc, t0, e0, F0 = call('uw12g_form_fock', h, s, kappa, ctx)
        assert rc == 0
        assert np.array_equal(t1, t0), (t1, t0)
        assert e1 == e0
        assert np.array_equal(F1, F0)
        # again, now with U^T Gamma kept on the handles by the allocating call
        rc, t2, e2, F2 = call('uw12g_form_fock_noalloc', h, s, kappa, ctx)
        assert rc == 0 and np.array_equal(t2, t0) and np.array_equal(F2, F0)


def test_noalloc_refuses_oversize_input():
    ctx = u.Context(0)
    s = synth.make_system('m12')
    h = gpu_setup(s, ctx)
    rc, *_ = call('uw12g_form_fock_noalloc', h, s, (1.0, 0.5), ctx)
    assert rc == -2
    msg = _lib.load().uw12g_last_error().decode()
    assert 'scratch' in msg, msg
    # the allocating entry point is unaffected, and a fitting call still works afterwards
    rc, *_ = call('uw12g_form_fock', h, s, (1.0, 0.5), ctx)
    assert rc == 0
    c1 = synth.make_system('c1')
    hc = gpu_setup(c1, ctx)
    rc, t1, _, F1 = call('uw12g_form_fock_noalloc', hc, c1, (1.0, 0.0), ctx)
    rc0, t0, _, F0 = call('uw12g_form_fock', hc, c1, (1.0, 0.0), ctx)
    assert rc == 0 and rc0 == 0 and np.array_equal(t1, t0) and np.array_equal(F1, F0)
