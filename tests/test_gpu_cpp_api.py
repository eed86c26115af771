"""The C++ mirror of libuw12.hpp (include/uw12g.hpp) driven from a compiled
C++ caller, checked against the CPU oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from helpers import fock_error, oracle_eval
from libuw12_amd import synth

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'cpp', 'uw12_cpp_driver')


def write(d, s, k0, k1):
    n_spin = len(s['orbitals'])
    n_orb = max(c.shape[1] for c in s['orbitals'])
    C = np.zeros((s['N'], n_orb * n_spin), order='F')
    for i, c in enumerate(s['orbitals']):
        C[:, i * n_orb:i * n_orb + c.shape[1]] = c
    occ = [len(x) for x in s['occ']] + [0]
    act = list(s['n_active']) + [0]
    with open(os.path.join(d, 'dims.txt'), 'w') as f:
        f.write(f"{s['N']} {s['Nd']} {s['Nr']} {n_orb} {n_spin} {occ[0]} {occ[1]} {act[0]} {act[1]} {k0!r} {k1!r}\n")
    arrays = dict(W3=s['ints']['W'][0], W2=s['ints']['W'][1], W3ri=s['ints']['W'][2], V3=s['ints']['V'][0],
                  V2=s['ints']['V'][1], V3ri=s['ints']['V'][2], WV3=s['ints']['WV'][0], WV2=s['ints']['WV'][1],
                  S=s['S'], C=C, occ=np.concatenate(s['occ']))
    for k, v in arrays.items():
        np.asfortranarray(v, dtype=np.float64).T.tofile(os.path.join(d, k + '.bin'))


@pytest.mark.parametrize('name', ['c1', 'm6u'])
def test_cpp_api_matches_oracle(tmp_path, name):
    if not os.path.exists(DRIVER):
        pytest.fail(f'{DRIVER} not built (python -m libuw12_amd.build)')
    s = synth.make_system(name)
    write(str(tmp_path), s, 1.0, 0.5)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    e1, e2 = (float(x) for x in open(tmp_path / 'energy.txt').read().split())
    n_spin = len(s['orbitals'])
    F = np.fromfile(tmp_path / 'fock.bin').reshape(n_spin * s['N'], s['N']).T
    Fs = [F[:, i * s['N']:(i + 1) * s['N']] for i in range(n_spin)]
    eo, Fo = oracle_eval(s, 1.0, 0.5)
    assert abs(e1 - eo['total']) <= 4e-9 and abs(e2 - eo['total']) <= 4e-9
    assert fock_error(Fs, Fo) <= 2e-8


def test_cpp_static_handles_outlive_context(tmp_path):
    """Handles kept in static storage are destroyed after the header's
    function-local context at process exit; the context stays alive until the
    last handle is gone, so the program exits cleanly."""
    if not os.path.exists(DRIVER):
        pytest.fail(f'{DRIVER} not built (python -m libuw12_amd.build)')
    s = synth.make_system('c1')
    write(str(tmp_path), s, 1.0, 0.5)
    r = subprocess.run([DRIVER, str(tmp_path), 'static'], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr)


def test_python_context_closed_before_handles(gpu_ctx):
    """Context.close() with live handles defers the context's deletion to the
    last handle's destroy."""
    import gc

    import libuw12_amd as u
    from helpers import gpu_eval, gpu_setup
    s = synth.make_system('c1')
    ctx = u.Context(0)
    h = gpu_setup(s, ctx)
    e_before, _ = gpu_eval(h, s, 1.0, 0.5)
    ctx.close()
    del h
    gc.collect()
    # the process's other contexts are unaffected
    e_after, _ = gpu_eval(gpu_setup(s, gpu_ctx), s, 1.0, 0.5)
    assert e_after['total'] == e_before['total']
