"""Writes tests/golden/c2_oracle.npz: the CPU oracle's energy pieces and Fock
matrix at the full c2 shape (n_ao=320, n_df=960, n_ri=1280, 80 occupied,
closed shell), for the GPU parity test tests/test_gpu_c2_parity.py.

Cases:
  recipe_k05   the s-Gaussian recipe (synth.make_system('c2')), kappa = (1, 0.5)
  recipe_k00   the same inputs, kappa = (1, 0)
  random_k05   well-conditioned random inputs of the same shape
               (synth.random_system(320, 960, 1280, [80], seed=2024), the style of
               the reference's unit-test fixtures), kappa = (1, 0.5)

Each case stores the six energy pieces (e2_direct ... e4_indirect), their total,
the Fock matrix, a SHA-256 digest of every input array and each array's sum and
sum of squares, so a change of the input generator shows up as a checksum
mismatch rather than as a parity failure.  The recipe's inputs are elementwise
formulas and reproduce bit for bit on any machine; the random case's metrics
come from BLAS products, so only its sums are compared (to 1e-12).

The oracle (oracle/uw12_oracle.cpp, the reference's contraction order in FP64)
takes a few minutes per case on 8 threads.  Run in the build container:
    python tests/golden/make_c2_fixtures.py
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'c2_oracle.npz')
TERMS = ('e2_direct', 'e2_indirect', 'e3_direct', 'e3_indirect', 'e4_direct', 'e4_indirect', 'total')


def digest(system):
    """SHA-256 of the inputs: integral arrays, S, orbitals, occupations (column-major bytes)."""
    h = hashlib.sha256()
    for kind in ('W', 'V', 'WV'):
        for a in system['ints'][kind]:
            if a is not None:
                h.update(np.asfortranarray(a, dtype=np.float64).tobytes(order='F'))
    h.update(np.asfortranarray(system['S'], dtype=np.float64).tobytes(order='F'))
    for c in system['orbitals']:
        h.update(np.asfortranarray(c, dtype=np.float64).tobytes(order='F'))
    for o in system['occ']:
        h.update(np.asarray(o, dtype=np.float64).tobytes())
    return h.hexdigest()


def input_norms(system):
    """Sum and sum of squares of every input array, in a fixed order: a check of the inputs
    that tolerates the last-bit differences BLAS gives across CPUs (the random case's
    metrics and overlap come from matrix products), where the SHA-256 digest does not."""
    arrays = [a for kind in ('W', 'V', 'WV') for a in system['ints'][kind] if a is not None]
    arrays += [system['S']] + list(system['orbitals']) + [np.asarray(o) for o in system['occ']]
    return np.array([[float(np.sum(a)), float(np.sum(np.square(a)))] for a in arrays])


def cases():
    from libuw12_amd import synth
    recipe = synth.make_system('c2')
    rnd = synth.random_system(320, 960, 1280, [80], seed=2024)
    return {'recipe_k05': (recipe, (1.0, 0.5)), 'recipe_k00': (recipe, (1.0, 0.0)), 'random_k05': (rnd, (1.0, 0.5))}


def main():
    from oracle import cpu
    out = {}
    for name, (system, (k0, k1)) in cases().items():
        t = time.perf_counter()
        sysc = cpu.System(system['ints'], system['S'])
        e, F = sysc.evaluate(system['orbitals'], system['occ'], system['n_active'], k0, k1, fock=True)
        print(f'{name}: {time.perf_counter() - t:.1f}s total {e["total"]:.12f}', flush=True)
        out[f'{name}_terms'] = np.array([e[k] for k in TERMS])
        out[f'{name}_fock'] = F[0]
        out[f'{name}_kappa'] = np.array([k0, k1])
        out[f'{name}_inputs_sha256'] = np.array(digest(system))
        out[f'{name}_inputs_norms'] = input_norms(system)
    out['terms'] = np.array(TERMS)
    np.savez_compressed(OUT, **out)
    print(f'wrote {OUT} ({os.path.getsize(OUT) / 1e6:.2f} MB)')


if __name__ == '__main__':
    main()
