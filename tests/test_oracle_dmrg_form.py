"""The oracle's DMRG executor (block-operator groups, super-block terms,
FiniteDMRG) against independent routes to the same numbers:

* UpdateLeftBlockOps / UpdateRightBlockOps (dmrg_init.h:131-205: per element,
  a sum of outer products, then two contractions) equal the MPO-form
  environment updates (vmps_init.h:343-373) sliced per bond element;
* the super-block terms of SetEffectiveHamiltonianTerms_ apply the same H_eff
  as the MPO-form matvec;
* finite_dmrg reaches the reference's FiniteDMRG known answers
  (test_finite_dmrg.cc) and the energies of the oracle's VMPS-form driver."""
import numpy as np
import pytest

from oracle.dmrg import SweepParams, UpdateSiteLenvs, UpdateSiteRenvs, direct_state_mps, init_envs, random_mps, \
    two_site_vmps
from oracle.lanczos import LanczosParams, eff_ham_mul_two_site_state
from oracle.models import build_mpo, ed_ground_energy, heisenberg_chain, hubbard_cylinder, hubbard_square
from oracle.superblock import (UpdateLeftBlockOps, UpdateRightBlockOps, _bond_elements, _id_scalar, finite_dmrg,
                               mat_repr_mpo, set_effective_hamiltonian_terms,
                               super_block_hamiltonian_mul_two_site_state)
from oracle.bsten import Contract


def _slice(env, e_ix, p):
    """Element p of an MPO-form environment env[x_bra, w, x_ket] as [x_ket, x_bra]."""
    s, q = _bond_elements(env.indexes[1])[p]
    out = {}
    for (x, ws, z), blk in env.blocks.items():
        if ws == s:
            out[(z, x)] = blk[:, q, :].T
    return out


def _close(ops, env, tol=1e-12):
    scale = max(float(np.abs(b).max()) for b in env.blocks.values())
    assert len(ops) == env.indexes[1].dim
    for p, op in enumerate(ops):
        ref = _slice(env, None, p)
        got = {} if op is None else op.blocks
        keys = set(ref) | set(got)
        for k in keys:
            a = got.get(k)
            b = ref.get(k)
            if a is None:
                assert np.abs(b).max() <= tol * scale
            elif b is None:
                assert np.abs(a).max() <= tol * scale
            else:
                assert np.abs(a - b).max() <= tol * scale


@pytest.mark.parametrize("which", ["heisenberg", "hubbard"])
def test_block_op_updates_equal_sliced_envs(which):
    if which == "heisenberg":
        m, n, total = heisenberg_chain(10), 10, (0,)
    else:
        m, n, total = hubbard_cylinder(3, 2), 6, (6, 0)
    mpo = build_mpo(m)
    mps = random_mps(m, total, 16, seed=7)
    mro = mat_repr_mpo(mpo)
    lenvs, renvs = init_envs(mps, mpo)
    # right groups from the right end (the MPS is right-canonical from site 1)
    rog = [_id_scalar(mps[-1].indexes[2], mps[0].dtype, False)]
    for i in range(1, n - 1):
        rog = UpdateRightBlockOps(rog, mps[n - i], mro[n - i])
        _close(rog, renvs[i])
    # left groups along a few sites
    log = [_id_scalar(mps[0].indexes[0], mps[0].dtype, True)]
    L = lenvs[0]
    for i in range(n // 2):
        log = UpdateLeftBlockOps(log, mps[i], mro[i])
        L = UpdateSiteLenvs(L, mps[i], mpo[i])
        _close(log, L)


def test_terms_apply_the_mpo_form_operator():
    m = hubbard_cylinder(3, 2)
    n = 6
    mpo = build_mpo(m)
    mps = random_mps(m, (6, 0), 16, seed=9)
    mro = mat_repr_mpo(mpo)
    lenvs, renvs = init_envs(mps, mpo)
    log = [_id_scalar(mps[0].indexes[0], mps[0].dtype, True)]
    for i in range(2):
        log = UpdateLeftBlockOps(log, mps[i], mro[i])
        lenvs[i + 1] = UpdateSiteLenvs(lenvs[i], mps[i], mpo[i])
    rog = [_id_scalar(mps[-1].indexes[2], mps[0].dtype, False)]
    for i in range(1, n - 3):
        rog = UpdateRightBlockOps(rog, mps[n - i], mro[n - i])
    l, r = 2, 3
    psi = Contract(mps[l], mps[r], ([2], [0]))
    terms = set_effective_hamiltonian_terms(log, rog, mro[l], mro[r])
    a = super_block_hamiltonian_mul_two_site_state(terms, psi)
    b = eff_ham_mul_two_site_state([lenvs[l], mpo[l], mpo[r], renvs[n - 1 - r]], psi)
    scale = max(float(np.abs(v).max()) for v in b.blocks.values())
    for k in set(a.blocks) | set(b.blocks):
        assert np.abs(a.blocks.get(k, 0) - b.blocks.get(k, 0)).max() <= 1e-12 * scale


def test_finite_dmrg_heisenberg_n6_kat():
    """test_finite_dmrg.cc:167-187 through the DMRG form: -2.493577133888 to 1e-12."""
    m = heisenberg_chain(6)
    mps = direct_state_mps(m, [i % 2 for i in range(6)])
    e = finite_dmrg(mps, mat_repr_mpo(build_mpo(m)), SweepParams(4, 8, 8, 1e-9, LanczosParams(1e-7)))
    assert abs(e - (-2.493577133888)) < 1e-12


def test_finite_dmrg_hubbard_2x2_kat():
    """test_finite_dmrg.cc:742-791 (t=1, t'=0.5, U=2): -2.828427124746 to 1e-10."""
    m = hubbard_square(2, 2, 1.0, 0.5, 2.0)
    mps = direct_state_mps(m, [1 if i % 2 == 0 else 2 for i in range(4)])
    e = finite_dmrg(mps, mat_repr_mpo(build_mpo(m)), SweepParams(10, 16, 16, 1e-9, LanczosParams(1e-8, 20)))
    assert abs(e - (-2.828427124746)) < 1e-10


def test_finite_dmrg_matches_vmps_form():
    """Same start, same parameters: every update's energy of the DMRG form
    equals the VMPS form's to 1e-10 (the two solve the same H_eff), and the
    end energy is the exact ground state of the L=8 chain."""
    m = heisenberg_chain(8)
    mpo = build_mpo(m)
    p = SweepParams(3, 16, 16, 1e-12, LanczosParams(1e-12, 60))
    log_d, log_v = [], []
    e_d = finite_dmrg(random_mps(m, (0,), 8, seed=5), mat_repr_mpo(mpo), p, log_d)
    e_v = two_site_vmps(random_mps(m, (0,), 8, seed=5), mpo, p, log_v)
    assert len(log_d) == len(log_v)
    for a, b in zip(log_d, log_v):
        assert abs(a["e0"] - b["e0"]) < 1e-10
    assert abs(e_d - e_v) < 1e-11
    assert abs(e_d - ed_ground_energy(m, (0,))) < 1e-9
