"""The oracle's DMRG-form (super-block) matvec equals its MPO-form matvec on the
same operators (the reference's two Lanczos overloads apply the same H_eff)."""
import numpy as np
import pytest

from oracle.lanczos import eff_ham_mul_two_site_state
from oracle.superblock import mpo_form_to_terms, super_block_hamiltonian_mul_two_site_state
from tests import refkat
from tests.workloads import synthetic_eff_ham


def _rel(a, b):
    keys = set(a.blocks) | set(b.blocks)
    num = max(float(np.abs(a.blocks.get(k, 0) - b.blocks.get(k, 0)).max()) for k in keys)
    den = max(float(np.abs(v).max()) for v in b.blocks.values())
    return num / den


@pytest.mark.parametrize("kind,D,seed", [("heisenberg", 24, 1), ("hubbard", 16, 2)])
def test_super_block_equals_mpo_form(kind, D, seed):
    eff, psi = synthetic_eff_ham(kind, D, seed=seed, nsect_half=2)
    terms = mpo_form_to_terms(*eff)
    assert terms, "no terms"
    a = super_block_hamiltonian_mul_two_site_state(terms, psi)
    b = eff_ham_mul_two_site_state(eff, psi)
    assert _rel(a, b) < 1e-12


def test_super_block_kat_operators():
    from oracle.bsten import BSTensor, random_fill
    L, Wl, Wr, R, (Din, dout, Dout) = refkat.kat_eff_ham(8, 2, 2)
    psi = random_fill(BSTensor([Din, dout, dout, Dout]), (0,), np.random.default_rng(3))
    a = super_block_hamiltonian_mul_two_site_state(mpo_form_to_terms(L, Wl, Wr, R), psi)
    b = eff_ham_mul_two_site_state([L, Wl, Wr, R], psi)
    assert _rel(a, b) < 1e-12
