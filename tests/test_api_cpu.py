"""Host-side argument handling of the Python mirror of libuw12.hpp (no GPU)."""
import numpy as np
import pytest

import libuw12_amd as u
from libuw12_amd import api


def test_setup_orbitals_splits_spin_blocks():
    C = np.arange(24.0).reshape(4, 6, order='F')
    a, b = u.setup_orbitals(C, 4, 3, 2)
    assert np.array_equal(a, C[:, :3]) and np.array_equal(b, C[:, 3:])
    (c,) = u.setup_orbitals(C, 4, 6, 1)
    assert np.array_equal(c, C)


def test_setup_orbitals_rejects_bad_spin():
    with pytest.raises(u.UW12Error):
        u.setup_orbitals(np.zeros((4, 9)), 4, 3, 3)


def test_setup_occupations():
    (o,) = u.setup_occupations([2.0, 2.0, 1.0], 3)
    assert list(o) == [2.0, 2.0, 1.0]
    a, b = u.setup_occupations([1.0, 1.0, 1.0, 0.5], 3, 1)
    assert list(a) == [1.0, 1.0, 1.0] and list(b) == [0.5]


def test_spin_channel_consistency_checked_before_device_work():
    C = [np.zeros((5, 3))]
    with pytest.raises(u.UW12Error):
        api._orbital_args(C, [np.ones(3), np.ones(2)], [3])
    with pytest.raises(u.UW12Error):
        api._orbital_args(C, [np.ones(3)], [3, 2])
    with pytest.raises(u.UW12Error):
        api._orbital_args([np.zeros((5, 3)), np.zeros((4, 3))], [np.ones(3), np.ones(3)], [3, 3])
    with pytest.raises(u.UW12Error):
        api._orbital_args(C, [np.ones(4)], [3])


def test_orbital_args_layout():
    C = [np.ones((5, 3)), 2 * np.ones((5, 2))]
    Cm, n_ao, n_orb, n_spin, occv, n_occ, n_act = api._orbital_args(C, [np.ones(3), np.ones(2)], [2, 1])
    assert (n_ao, n_orb, n_spin) == (5, 3, 2)
    assert Cm.shape == (5, 6) and np.all(Cm[:, :3] == 1) and np.all(Cm[:, 3:5] == 2) and np.all(Cm[:, 5] == 0)
    assert list(n_occ) == [3, 2] and list(n_act) == [2, 1]
