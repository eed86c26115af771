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


def density_derivative(F, orbitals, occ, dC):
    """sum_s <F_s, dD_s/dt> at t = 0 for D_s(t) = (C_s + t dC_s) diag(occ_s) (C_s + t dC_s)^T."""
    tot = 0.0
    for s in range(len(orbitals)):
        o = len(occ[s])
        C = np.asarray(orbitals[s], dtype=np.float64)[:, :o]
        d = np.asarray(dC[s], dtype=np.float64)[:, :o]
        wC = C * np.asarray(occ[s], dtype=np.float64)
        tot += float(np.sum(F[s] * (d @ wC.T + wC @ d.T)))
    return tot


def active_only(dC, occ, n_active):
    """The directions with their core columns (the first n_occ - n_active occupied) zeroed:
    paths that move only the active orbitals."""
    out = []
    for s, d in enumerate(dC):
        d = np.array(d, dtype=np.float64)
        d[:, :len(occ[s]) - int(n_active[s])] = 0.0
        out.append(d)
    return out


def ld_directional_derivative(system, k0, k1, dC, n_active=None, h=5e-5, orbitals=None, occ=None):
    """dE/dt at t = 0 along C(t) = C + t dC from long-double energies (oracle/extended.py), by
    Richardson-extrapolated central differences: every product and sum in 64-bit-significand
    arithmetic.  On the c1 recipe the 3e / 4e pieces of dE/dt are ~100 against a total of ~1,
    so the O(h^4) truncation is 4e-7 relative at h = 1e-3 and 3e-9 at 3e-4; at the default
    5e-5 it is ~1e-12, with ~1e-13 of rounding."""
    from oracle import extended
    LD = np.longdouble
    C0 = [np.asarray(c).astype(LD) for c in (orbitals or system['orbitals'])]
    D = [np.asarray(d).astype(LD) for d in dC]
    occ = occ or system['occ']
    act = n_active or system['n_active']

    def E(t):
        C = [c + LD(t) * d for c, d in zip(C0, D)]
        return extended.energy_terms(system['ints'], system['S'], C, occ, act, k0, k1, raw=True)['total']

    d1 = (E(h) - E(-h)) / LD(2 * h)
    d2 = (E(2 * h) - E(-2 * h)) / LD(4 * h)
    return float((LD(4) * d1 - d2) / LD(3))
