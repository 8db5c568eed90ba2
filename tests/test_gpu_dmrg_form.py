"""The DMRG-form executor on the GPU (ultradmrg_amd/dmrg.py) against the
oracle's restatement of dmrg_impl.h / dmrg_init.h and the reference's
FiniteDMRG known answers (test_finite_dmrg.cc).

Bars: the block-sparse transpose is bit-exact against numpy; block-operator
groups (every element) and the super-block matvec built from
SetEffectiveHamiltonianTerms_ match the oracle to 1e-12 relative; FiniteDMRG
energies match the reference's known answers to the reference's own
tolerances and the VMPS-form GPU driver update by update to 1e-10."""
import numpy as np
import pytest
import torch

from oracle.bsten import BSTensor, Contract, Index, random_fill
from oracle.dmrg import init_envs, random_mps as o_random_mps
from oracle.models import build_mpo as o_build_mpo, heisenberg_chain as o_heis, hubbard_cylinder as o_hub
from oracle.superblock import (UpdateLeftBlockOps as o_ulbo, UpdateRightBlockOps as o_urbo, _id_scalar,
                               mat_repr_mpo as o_mat_repr, set_effective_hamiltonian_terms,
                               super_block_hamiltonian_mul_two_site_state)
from tests.convert import from_device, to_device, to_oracle

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd import mpo as M
    from ultradmrg_amd.dmrg import (DMRGExecutor, FiniteDMRG, UpdateLeftBlockOps, UpdateRightBlockOps,
                                    mat_repr_mpo)
    from ultradmrg_amd.solver import EffectiveHamiltonian, LanczosParams
    from ultradmrg_amd.sweep import SweepParams, TwoSiteVMPS, direct_state_mps, random_mps
    from ultradmrg_amd.tensor import transpose


def _rand_tensor(rng, rank, cplx):
    qs = [((q,), int(rng.integers(1, 4))) for q in range(-1, 2)]
    dirs = [(-1) ** int(rng.integers(0, 2)) for _ in range(rank)]
    t = BSTensor([Index(tuple(qs), d) for d in dirs], np.complex128 if cplx else np.float64)
    return random_fill(t, (0,), rng, complex_=cplx)


@pytest.mark.parametrize("rank,cplx", [(2, False), (3, False), (3, True), (4, False), (4, True)])
def test_transpose_bit_exact(rank, cplx):
    print(f"[case] transpose rank {rank} complex {cplx}", flush=True)
    rng = np.random.default_rng(rank * 10 + cplx)
    t = _rand_tensor(rng, rank, cplx)
    d = to_device(t)
    for _ in range(4):
        perm = list(rng.permutation(rank))
        ref = t.transpose(perm)
        got = transpose(d, perm)
        assert got.coords == sorted(ref.blocks)
        hb = got.blocks()
        for k, v in ref.blocks.items():
            assert np.array_equal(hb[k], v)


def _groups_problem(which):
    if which == "heisenberg":
        m, n, total = o_heis(10), 10, (0,)
    else:
        m, n, total = o_hub(3, 2), 6, (6, 0)
    return m, n, o_build_mpo(m), o_random_mps(m, total, 24, seed=41)


def _check_group(group, ref_ops, tol=1e-12):
    ops = group.ops()
    assert len(ops) == len(ref_ops)
    scale = max(float(np.abs(b).max()) for r in ref_ops if r is not None for b in r.blocks.values())
    for got, ref in zip(ops, ref_ops):
        gb = {} if got is None else got.blocks()
        rb = {} if ref is None else ref.blocks
        for k in set(gb) | set(rb):
            a, b = gb.get(k), rb.get(k)
            diff = np.abs(a if b is None else (b if a is None else a - b)).max()
            assert diff <= tol * scale, (k, diff, scale)


@pytest.mark.parametrize("which", ["heisenberg", "hubbard"])
def test_block_operator_groups_parity(which):
    """DMRGInit_ right groups of every length and left groups over half the
    chain vs the oracle's per-element updates (dmrg_init.h:131-205)."""
    print(f"[case] block operator groups {which}", flush=True)
    m, n, mpo, mps = _groups_problem(which)
    omro = o_mat_repr(mpo)
    dmps = [to_device(t) for t in mps]
    ex = DMRGExecutor(dmps, mat_repr_mpo([to_device(w) for w in mpo]),
                      SweepParams(1, 24, 24, 1e-10, LanczosParams(1e-10, 50), verbose=False))
    rog = [_id_scalar(mps[-1].indexes[2], mps[0].dtype, False)]
    for i in range(1, n - 1):
        rog = o_urbo(rog, mps[n - i], omro[n - i])
        _check_group(ex.ropg[i], rog)
    log = [_id_scalar(mps[0].indexes[0], mps[0].dtype, True)]
    g = ex.lopg[0]
    for i in range(n // 2):
        log = o_ulbo(log, mps[i], omro[i])
        g = UpdateLeftBlockOps(g, dmps[i], ex.W[i], ex.bonds[i + 1])
        _check_group(g, log)


@pytest.mark.parametrize("which", ["heisenberg", "hubbard"])
def test_super_block_terms_matvec_parity(which):
    """SetEffectiveHamiltonianTerms_ on the GPU groups, planned as a
    super-block operator, vs the oracle's term-wise matvec on the oracle's
    groups (lanczos_dmrg_solver_impl.h:199-283)."""
    print(f"[case] super-block terms matvec {which}", flush=True)
    m, n, mpo, mps = _groups_problem(which)
    omro = o_mat_repr(mpo)
    dmps = [to_device(t) for t in mps]
    ex = DMRGExecutor(dmps, mat_repr_mpo([to_device(w) for w in mpo]),
                      SweepParams(1, 24, 24, 1e-10, LanczosParams(1e-10, 50), verbose=False))
    l, r = 0, 1
    log = [_id_scalar(mps[0].indexes[0], mps[0].dtype, True)]
    rog = [_id_scalar(mps[-1].indexes[2], mps[0].dtype, False)]
    for i in range(1, n - 1 - l):
        rog = o_urbo(rog, mps[n - i], omro[n - i])
    psi = Contract(mps[l], mps[r], ([2], [0]))
    ref = super_block_hamiltonian_mul_two_site_state(set_effective_hamiltonian_terms(log, rog, omro[l], omro[r]), psi)
    terms = ex.SetEffectiveHamiltonianTerms_(l, r)
    dpsi = to_device(psi)
    H = EffectiveHamiltonian(terms, dpsi)
    out = from_device(H @ dpsi)
    H.close()
    scale = max(float(np.abs(v).max()) for v in ref.blocks.values())
    for k in set(out.blocks) | set(ref.blocks):
        a, b = out.blocks.get(k), ref.blocks.get(k)
        diff = np.abs(a if b is None else (b if a is None else a - b)).max()
        assert diff <= 1e-12 * scale


def test_finite_dmrg_heisenberg_n6_kat():
    """test_finite_dmrg.cc:167-187 (FiniteDMRG): -2.493577133888 to 1e-12."""
    print("[case] FiniteDMRG heisenberg N=6", flush=True)
    m = M.heisenberg_chain(6)
    mps = direct_state_mps(m, [i % 2 for i in range(6)])
    e, _ = FiniteDMRG(mps, mat_repr_mpo(M.build_mpo(m)), SweepParams(4, 8, 8, 1e-9, LanczosParams(1e-7),
                                                                      verbose=False))
    assert abs(e - (-2.493577133888)) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
def test_finite_dmrg_heisenberg_2d_kat(cplx):
    """test_finite_dmrg.cc:223-281 (FiniteDMRG, real and complex): -3.129385241572 to 1e-12."""
    print(f"[case] FiniteDMRG 2D heisenberg complex={cplx}", flush=True)
    dt = torch.complex128 if cplx else torch.float64
    m = M.Model(6, M.SPIN_QNS, M.spin_operators())
    for i, j in [(0, 1), (0, 2), (1, 3), (2, 3), (2, 4), (3, 5), (4, 5)]:
        m.add(1.0, (i, "sz"), (j, "sz"))
        m.add(0.5, (i, "sp"), (j, "sm"))
        m.add(0.5, (i, "sm"), (j, "sp"))
    mps = direct_state_mps(m, [i % 2 for i in range(6)], dtype=dt)
    e, _ = FiniteDMRG(mps, mat_repr_mpo(M.build_mpo(m, dtype=dt)),
                      SweepParams(4, 8, 8, 1e-9, LanczosParams(1e-7), verbose=False))
    assert abs(e - (-3.129385241572)) < 1e-12


def test_finite_dmrg_hubbard_2x2_kat():
    """test_finite_dmrg.cc:742-791 (FiniteDMRG, t=1, t'=0.5, U=2): -2.828427124746 to 1e-10."""
    print("[case] FiniteDMRG hubbard 2x2", flush=True)
    m = M.Model(4, M.HUBBARD_QNS, M.hubbard_operators())
    idx = lambda x, y: x * 2 + y
    for i in range(2):
        for j in range(2):
            s0 = idx(i, j)
            m.add(2.0, (s0, "Uterm"))
            if i != 1:
                M._hopping(m, 1.0, s0, idx(i + 1, j))
            if j != 1:
                M._hopping(m, 1.0, s0, idx(i, j + 1))
                if i != 0:
                    M._hopping(m, 0.5, s0, idx(i - 1, j + 1))
                if i != 1:
                    M._hopping(m, 0.5, s0, idx(i + 1, j + 1))
    mps = direct_state_mps(m, [1 if i % 2 == 0 else 2 for i in range(4)])
    e, _ = FiniteDMRG(mps, mat_repr_mpo(M.build_mpo(m)),
                      SweepParams(10, 16, 16, 1e-9, LanczosParams(1e-8, 20), verbose=False))
    assert abs(e - (-2.828427124746)) < 1e-10


def test_finite_dmrg_heisenberg_l20_matches_vmps_form():
    """Open Heisenberg chain L=20, Sz=0, seeded random MPS at D=64, 4 sweeps,
    Lanczos(1e-10, 100): the DMRG form reaches the exact
    -8.682473334398967 (1e-10 relative) and agrees with the VMPS form update
    by update to 1e-10."""
    print("[case] FiniteDMRG L=20 D=64 vs VMPS", flush=True)
    m = M.heisenberg_chain(20)
    p = SweepParams(4, 64, 64, 1e-10, LanczosParams(1e-10, 100), verbose=False)
    e_d, ex = FiniteDMRG(random_mps(m, (0,), 64, seed=1001), mat_repr_mpo(M.build_mpo(m)), p)
    drv = TwoSiteVMPS(random_mps(m, (0,), 64, seed=1001), M.build_mpo(m), p)
    e_v = drv.run()
    assert abs(e_d - (-8.682473334398967)) < 1e-10 * 8.68
    assert len(ex.stats.log) == len(drv.stats.log)
    worst = max(abs(a["e0"] - b["e0"]) for a, b in zip(ex.stats.log, drv.stats.log))
    print(f"  E DMRG form {e_d:.15f} VMPS form {e_v:.15f} worst per-update difference {worst:.2e}; "
          f"sweep s DMRG {[round(s, 3) for s in ex.stats.sweep_seconds]} VMPS "
          f"{[round(s, 3) for s in drv.stats.sweep_seconds]}", flush=True)
    assert worst < 1e-10
