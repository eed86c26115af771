"""Packs the reference's water test data (zwill22/libuw12 test_data/water/*.csv,
read as plain text) into tests/golden/water.npz.

W3_ri.csv and V3_ri.csv are absent from the reference checkout, so the golden
energies (-0.344173824 / -0.272864310 Eh, tests/test_water.cpp:12-78) cannot be
reproduced; the arrays serve as realistic molecular DF inputs (n_ao=24,
n_df=116, n_ri=76) for the RI-free two- and four-electron terms.
Run in the build container: python tests/golden/make_water_fixture.py
"""
import os

import numpy as np

SRC = '/root/reference/test_data/water'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'water.npz')


def load(name):
    return np.loadtxt(os.path.join(SRC, name + '.csv'), delimiter=',', ndmin=2)


def main():
    arrays = {k: load(k) for k in ('W2', 'W3', 'V2', 'V3', 'WV2', 'WV3', 'S', 'orbitals', 'orbitals_alpha',
                                   'orbitals_beta', 'occ', 'occ_alpha', 'occ_beta', 'fock')}
    arrays['n_active_neutral'] = np.array([4])
    arrays['n_active_cation'] = np.array([4, 3])
    arrays['energy_neutral'] = np.array([-0.344173824])
    arrays['energy_cation'] = np.array([-0.272864310])
    np.savez_compressed(OUT, **arrays)
    print({k: v.shape for k, v in arrays.items()})


if __name__ == '__main__':
    main()
