"""Differentiation form at a full configuration: the adjoint Fock against
central finite differences of the energy.

Usage (on a GPU): python -m libuw12_amd.fd_check [--config c4] [--kappa 1.0,0.0] [--dirs 16]

For each of `dirs` seeded directions the occupied orbitals move along an
S-consistent path C(t) = C + t dC, dC = (1 - C C^T S) G, G ~ N(0, 1) 1e-4, so
the density D(t) = g Co(t) Co(t)^T stays in the occupied space's tangent
directions.  The FP64 energies at t = +-1 and t = +-2 give the central
differences d1 = (E(1) - E(-1)) / 2 and d2 = (E(2) - E(-2)) / 4 and their
Richardson extrapolation (4 d1 - d2) / 3, whose truncation error is O(t^4).
That is compared with the directional derivative sum_ab F_ab dD_ab of the
adjoint Fock (uw12g_forward + uw12g_adjoint), for the FP64 and the mixed
form; errors are given relative to the largest |dE/dt| over the directions.
The adjoint Fock is also compared with the fused FP64 Fock (uw12g_form_fock).
The larger configurations stream their integrals per DF shell, as bench.py
does.  Prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--config', default='c4', choices=['c1', 'c2', 'c3', 'c4', 'm6', 'm6u', 'm12'])
    ap.add_argument('--kappa', default='1.0,0.0')
    ap.add_argument('--dirs', type=int, default=16)
    ap.add_argument('--scale', type=float, default=1e-4)
    args = ap.parse_args()
    k0, k1 = (float(x) for x in args.kappa.split(','))

    import libuw12_amd as u
    from libuw12_amd import synth

    stream = args.config in ('c3', 'c4')
    t = time.perf_counter()
    s = synth.make_system(args.config, with_integrals=not stream)
    N, Nr = s['N'], s['Nr']
    ctx = u.Context(0)
    if stream:
        hs = []
        for kind in ('W', 'V', 'WV'):
            X2, sizes, fn, fn_ri = synth.shell_integrals(s, kind)
            hs.append(u.setup_base_integrals_fn(X2, fn, fn_ri, sizes, N, Nr if fn_ri else 0, context=ctx))
        W, V, WV = hs
    else:
        W = u.setup_base_integrals(*s['ints']['W'], n_ao=N, context=ctx)
        V = u.setup_base_integrals(*s['ints']['V'], n_ao=N, context=ctx)
        WV = u.setup_base_integrals(s['ints']['WV'][0], s['ints']['WV'][1], None, n_ao=N, context=ctx)
    P = u.setup_abs_projectors(s['S'], N, Nr, context=ctx)
    print(f'[fd_check] {args.config} set up in {time.perf_counter() - t:.1f}s', file=sys.stderr, flush=True)
    orbs, occ, act = s['orbitals'], s['occ'], s['n_active']
    ind = k1 != 0.0
    n_spin = len(orbs)
    g = 2.0 if n_spin == 1 else 1.0
    kw = dict(indirect_term=ind, scale_opp_spin=k0, scale_same_spin=k1)

    F_adj = {}
    for name, prec in (('fp64', u.FP64), ('mixed', u.MIXED)):
        t = time.perf_counter()
        _, tape = u.forward(W, V, WV, P, orbs, occ, act, **kw, precision=prec)
        print(f'[fd_check] {name} forward {time.perf_counter() - t:.1f}s', file=sys.stderr, flush=True)
        t = time.perf_counter()
        F_adj[name] = u.adjoint(tape)
        print(f'[fd_check] {name} adjoint {time.perf_counter() - t:.1f}s', file=sys.stderr, flush=True)
    t = time.perf_counter()
    _, F_fused = u.form_fock(W, V, WV, P, orbs, occ, act, **kw, precision=u.FP64)
    print(f'[fd_check] fused FP64 Fock {time.perf_counter() - t:.1f}s', file=sys.stderr, flush=True)

    Sao = s['S'][:N, :N]
    rng = np.random.default_rng(4004)
    rows = []
    for d in range(args.dirs):
        dC = []
        for sp in range(n_spin):
            C = orbs[sp][:, :len(occ[sp])]
            G = rng.standard_normal(C.shape) * args.scale
            full = np.zeros_like(orbs[sp])
            full[:, :C.shape[1]] = G - C @ (C.T @ (Sao @ G))
            dC.append(full)
        E = {}
        for tt in (1.0, -1.0, 2.0, -2.0):
            Cs = [np.asfortranarray(c + tt * x) for c, x in zip(orbs, dC)]
            e, _ = u.form_fock(W, V, WV, P, Cs, occ, act, calculate_fock=False, **kw, precision=u.FP64)
            E[tt] = e['total']
            print(f'[fd_check] direction {d + 1}: energy at t = {tt:+.0f}', file=sys.stderr, flush=True)
        d1 = 0.5 * (E[1.0] - E[-1.0])
        d2 = 0.25 * (E[2.0] - E[-2.0])
        fd = (4.0 * d1 - d2) / 3.0
        an = {}
        for name, F in F_adj.items():
            v = 0.0
            for sp in range(n_spin):
                o = len(occ[sp])
                w = np.sqrt(np.asarray(occ[sp]) / g)
                Co, dCo = orbs[sp][:, :o] * w, dC[sp][:, :o] * w
                v += np.sum(F[sp] * g * (dCo @ Co.T + Co @ dCo.T))
            an[name] = v
        rows.append({'fd': fd, 'fd_central': d1, **{f'{k}_err': v - fd for k, v in an.items()}})
        print(f'[fd_check] direction {d + 1}/{args.dirs}: fd {fd:.6e}, adjoint fp64 {an["fp64"]:.6e}, '
              f'mixed {an["mixed"]:.6e}', file=sys.stderr, flush=True)
    scale = max(abs(r['fd']) for r in rows)
    out = {
        'config': args.config, 'kappa': [k0, k1], 'dirs': args.dirs, 'scale': args.scale,
        'max_abs_fd': scale,
        'max_err_rel_to_max_fd': {k: max(abs(r[f'{k}_err']) for r in rows) / scale for k in F_adj},
        'central_vs_richardson_rel_to_max_fd': max(abs(r['fd_central'] - r['fd']) for r in rows) / scale,
        'adjoint_vs_fused_fp64_max_abs': {k: max(float(np.abs(a - b).max()) for a, b in zip(F, F_fused))
                                          for k, F in F_adj.items()},
        'fock_max_abs': max(float(np.abs(b).max()) for b in F_fused),
        'directions': rows,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
