"""Shared drivers for the parity tests: the same inputs through the GPU library
(C ABI) and through the CPU oracle."""
import numpy as np

TERMS = ('e2_direct', 'e2_indirect', 'e3_direct', 'e3_indirect', 'e4_direct', 'e4_indirect')


def gpu_setup(system, ctx):
    import libuw12_amd as u
    N, Nr = system['N'], system['Nr']
    ints = system['ints']
    W = u.setup_base_integrals(*ints['W'], n_ao=N, context=ctx)
    V = u.setup_base_integrals(*ints['V'], n_ao=N, context=ctx)
    WV = u.setup_base_integrals(ints['WV'][0], ints['WV'][1], None, n_ao=N, context=ctx)
    P = u.setup_abs_projectors(system['S'], N, Nr, context=ctx)
    return W, V, WV, P


def gpu_eval(handles, system, k0, k1, fock=True, indirect=None, orbitals=None, occ=None, n_active=None):
    import libuw12_amd as u
    W, V, WV, P = handles
    if indirect is None:
        indirect = abs(k1) > 4 * np.finfo(float).eps
    return u.form_fock(W, V, WV, P, orbitals or system['orbitals'], occ or system['occ'],
                       n_active or system['n_active'], indirect_term=indirect, calculate_fock=fock,
                       scale_opp_spin=k0, scale_same_spin=k1)


def oracle_eval(system, k0, k1, fock=True, indirect=None, orbitals=None, occ=None, n_active=None):
    from oracle import cpu
    return cpu.uw12(system['ints'], system['S'], orbitals or system['orbitals'], occ or system['occ'],
                    n_active or system['n_active'], k0, k1, indirect=indirect, fock=fock)


def term_errors(a, b):
    return {k: abs(a[k] - b[k]) for k in TERMS + ('total',)}


def fock_error(Fa, Fb):
    return max(float(np.abs(x - y).max()) for x, y in zip(Fa, Fb))


def fock_scale(F):
    return max(float(np.abs(x).max()) for x in F)
