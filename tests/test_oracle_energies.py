"""End-to-end energies of the oracle's two-site DMRG vs the reference's known answers."""
import numpy as np
import pytest

from oracle.dmrg import SweepParams, direct_state_mps, two_site_vmps, random_mps
from oracle.lanczos import LanczosParams
from oracle.models import build_mpo, ed_ground_energy, heisenberg_chain, hubbard_square, hubbard_cylinder


def test_heisenberg_n6():
    """tests/test_algorithm/test_finite_dmrg.cc:167-187: -2.493577133888 to 1e-12."""
    m = heisenberg_chain(6)
    mps = direct_state_mps(m, [i % 2 for i in range(6)])
    e = two_site_vmps(mps, build_mpo(m), SweepParams(4, 8, 8, 1e-9, LanczosParams(1e-7)))
    assert abs(e - (-2.493577133888)) < 1e-12
    assert abs(ed_ground_energy(m, (0,)) - (-2.493577133888)) < 1e-12


def test_hubbard_2x2():
    """test_finite_dmrg.cc:729-791 (t=1, t'=0.5, U=2, U1xU1): -2.828427124746 to 1e-10."""
    m = hubbard_square(2, 2, 1.0, 0.5, 2.0)
    mps = direct_state_mps(m, [1 if i % 2 == 0 else 2 for i in range(4)])
    e = two_site_vmps(mps, build_mpo(m), SweepParams(10, 16, 16, 1e-9, LanczosParams(1e-8, 20)))
    assert abs(e - (-2.828427124746)) < 1e-10
    assert abs(ed_ground_energy(m, (4, 0)) - (-2.828427124746)) < 1e-10


def test_hubbard_cylinder_mpo_vs_ed():
    """Snake 3x2 cylinder-like lattice (Ly=2 keeps it ED-sized): DMRG energy equals ED."""
    m = hubbard_cylinder(3, 2, t=1.0, U=8.0)
    e_ed = ed_ground_energy(m, (6, 0))
    labs = [1 if i % 2 == 0 else 2 for i in range(6)]
    mps = direct_state_mps(m, labs)
    e = two_site_vmps(mps, build_mpo(m), SweepParams(6, 64, 64, 1e-12, LanczosParams(1e-11, 100)))
    assert abs(e - e_ed) < 1e-9


def test_random_mps_start():
    m = heisenberg_chain(10)
    mps = random_mps(m, (0,), 16, seed=1001)
    e = two_site_vmps(mps, build_mpo(m), SweepParams(3, 32, 32, 1e-12, LanczosParams(1e-10, 100)))
    assert abs(e - ed_ground_energy(m, (0,))) < 1e-9
