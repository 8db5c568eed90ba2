"""End-to-end energies of the oracle's two-site DMRG vs the reference's known answers."""
import numpy as np
import pytest

from oracle.dmrg import SweepParams, direct_state_mps, two_site_vmps, random_mps
from oracle.lanczos import LanczosParams
from oracle.models import (HUBBARD_QNS, SPIN_QNS, ModelSpec, _add_hopping, build_mpo, ed_ground_energy, heisenberg_chain,
                           hubbard_ops, hubbard_square, hubbard_cylinder, spin_ops)

# test_finite_dmrg.cc:209-218: the 6-site 2D Heisenberg bonds
HEIS_2D_BONDS = [(0, 1), (0, 2), (1, 3), (2, 3), (2, 4), (3, 5), (4, 5)]


def ising_chain(n):
    """test_finite_dmrg.cc:109-113: H = sum Sz_i Sz_{i+1}."""
    m = ModelSpec(n, SPIN_QNS, spin_ops())
    for i in range(n - 1):
        m.add(1.0, (i, "sz"), (i + 1, "sz"))
    return m


def heisenberg_bonds(n, bonds):
    m = ModelSpec(n, SPIN_QNS, spin_ops())
    for i, j in bonds:
        m.add(1.0, (i, "sz"), (j, "sz"))
        m.add(0.5, (i, "sp"), (j, "sm"))
        m.add(0.5, (i, "sm"), (j, "sp"))
    return m


def free_fermion_chain(n):
    """Open chain H = -sum (c+_i c_{i+1} + h.c.) for both spins (Hubbard at U=0);
    in the sector of spin-up electrons only it is the spinless free-fermion chain
    of test_two_site_finite_tdvp.cc:102-106."""
    m = ModelSpec(n, HUBBARD_QNS, hubbard_ops())
    for i in range(n - 1):
        _add_hopping(m, 1.0, i, i + 1)
    return m


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


def test_ising_n6():
    """test_finite_dmrg.cc:109-127: -0.25 (N - 1) to 1e-10 (4 sweeps, D 1..10, LanczosParams(1e-7))."""
    m = ising_chain(6)
    mps = direct_state_mps(m, [i % 2 for i in range(6)])
    e = two_site_vmps(mps, build_mpo(m), SweepParams(4, 1, 10, 1e-5, LanczosParams(1e-7)))
    assert abs(e - (-0.25 * 5)) < 1e-10
    assert abs(ed_ground_energy(m, (0,)) - (-0.25 * 5)) < 1e-12


def test_heisenberg_2d_6site():
    """test_finite_dmrg.cc:207-254: -3.129385241572 to 1e-12 (4 sweeps, D=8, LanczosParams(1e-7))."""
    m = heisenberg_bonds(6, HEIS_2D_BONDS)
    mps = direct_state_mps(m, [i % 2 for i in range(6)])
    e = two_site_vmps(mps, build_mpo(m), SweepParams(4, 8, 8, 1e-9, LanczosParams(1e-7)))
    assert abs(e - (-3.129385241572)) < 1e-12
    assert abs(ed_ground_energy(m, (0,)) - (-3.129385241572)) < 1e-12


def test_free_fermion_n6_energy():
    """test_two_site_finite_tdvp.cc:121-123: -3.4939592074349334893668128643185 to 1e-13, here as three
    spin-up electrons of the U=0 Hubbard chain (N=3, 2Sz=3); the closed form is
    -2 (cos(pi/7) + cos(2 pi/7) + cos(3 pi/7))."""
    ref = -3.4939592074349334893668128643185
    assert abs(-2 * sum(np.cos(k * np.pi / 7) for k in (1, 2, 3)) - ref) < 1e-14
    m = free_fermion_chain(6)
    assert abs(ed_ground_energy(m, (3, 3)) - ref) < 1e-13
    mps = direct_state_mps(m, [1, 3, 1, 3, 1, 3])
    e = two_site_vmps(mps, build_mpo(m), SweepParams(4, 1, 16, 1e-10, LanczosParams(1e-8)))
    assert abs(e - ref) < 1e-13


@pytest.mark.parametrize("case", ["1d", "2d"])
def test_tj(case):
    """test_finite_dmrg.cc:484-583 (t=3, J=1, N=4, two electrons, Sz=0): 1D -6.947478526233 (11 sweeps),
    2D plaquette -8.868563739680 (10 sweeps), D=8, LanczosParams(1e-8, 20), to 1e-10."""
    from tests.workloads import TJ_1D_BONDS, TJ_2D_BONDS, tj_model
    bonds, labs, sweeps, ref = ((TJ_1D_BONDS, [2, 1, 2, 0], 11, -6.947478526233) if case == "1d" else
                                (TJ_2D_BONDS, [2, 0, 1, 2], 10, -8.868563739680))
    m = tj_model(ModelSpec, 4, bonds, 3.0, 1.0)
    assert abs(ed_ground_energy(m, (2, 0)) - ref) < 1e-10
    e = two_site_vmps(direct_state_mps(m, labs), build_mpo(m), SweepParams(sweeps, 8, 8, 1e-9, LanczosParams(1e-8, 20)))
    assert abs(e - ref) < 1e-10


def test_rashba_exact():
    """test_finite_dmrg.cc:620-706: the Rashba t-J model (charge-only U(1), complex spin-orbit
    coupling, 3x2 lattice periodic in x) has E0 = -11.018692166942165 at N=5: dense exact
    diagonalisation of the model as transcribed for the GPU test."""
    import itertools
    from tests.workloads import RASHBA_STD_OPS as O, rashba_terms
    n = 6

    def op_at(d):
        m = np.array([[1.0 + 0j]])
        for k in range(n):
            m = np.kron(m, d.get(k, np.eye(3)))
        return m
    H = np.zeros((3 ** n, 3 ** n), complex)
    for c, oa, a, ob, b, string in rashba_terms():
        d = {k: O["f"] for k in range(a + 1, b)} if string else {}
        d[a], d[b] = O[oa], O[ob]
        H += c * op_at(d)
    assert np.abs(H - H.conj().T).max() < 1e-14
    N = np.array([sum(1 for s in cfg if s < 2) for cfg in itertools.product(range(3), repeat=n)])
    keep = np.where(N == 5)[0]
    assert abs(np.linalg.eigvalsh(H[np.ix_(keep, keep)])[0] - (-11.018692166942165)) < 1e-12


def test_kondo_exact():
    """test_finite_dmrg.cc:846-952: the Kondo double chain (extended electrons on even sites, localized
    spins on odd sites, t=0.25, Jk=1, Jz=0.5, 8 sites) has E0 = -3.180025784229132: dense exact
    diagonalisation of the model as transcribed for the GPU test (4096 states)."""
    from tests.workloads import KONDO_STD_OPS as O, kondo_terms
    n = 8
    dims = [4 if i % 2 == 0 else 2 for i in range(n)]

    def op_at(d):
        m = np.array([[1.0]])
        for k in range(n):
            m = np.kron(m, d.get(k, np.eye(dims[k])))
        return m
    H = sum(c * op_at({a: O[oa], b: O[ob]}) for c, oa, a, ob, b in kondo_terms())
    assert np.abs(H - H.T).max() == 0.0
    assert abs(np.linalg.eigvalsh(H)[0] - (-3.180025784229132)) < 1e-12
