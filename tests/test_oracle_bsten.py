"""Block-sparse oracle vs dense brute force (the dense-contraction identity)."""
import numpy as np
import pytest

from oracle.bsten import BSTensor, Contract, ContractRotated, Dag, LinearCombine, vdot
from oracle.dmrg import UpdateSiteLenvs, UpdateSiteRenvs, random_mps, init_envs
from oracle.lanczos import eff_ham_mul_two_site_state, eff_ham_mul_single_site_state, two_site_matvec_flops
from oracle.models import build_mpo, heisenberg_chain, hubbard_cylinder
from tests.workloads import synthetic_eff_ham


@pytest.mark.parametrize("kind,D", [("heisenberg", 24), ("hubbard", 16)])
def test_two_site_matvec_vs_dense(kind, D):
    eff, psi = synthetic_eff_ham(kind, D, seed=7, nsect_half=3 if kind == "heisenberg" else 1, width=1.5)
    L, Wl, Wr, R = eff
    res = eff_ham_mul_two_site_state(eff, psi)
    # result carries the state's own index order
    assert [ix.sectors for ix in res.indexes] == [ix.sectors for ix in psi.indexes]
    assert [ix.dir for ix in res.indexes] == [ix.dir for ix in psi.indexes]
    t = np.einsum("xwa,asuc->xwsuc", L.to_dense(), psi.to_dense())
    t = np.einsum("xwsuc,wsty->xtyuc", t, Wl.to_dense())
    t = np.einsum("xtyuc,yuvz->xtvzc", t, Wr.to_dense())
    dense = np.einsum("xtvzc,bzc->xtvb", t, R.to_dense())
    assert np.allclose(res.to_dense(), dense, rtol=1e-12, atol=1e-12 * np.abs(dense).max())
    # every produced block obeys the state's charge rule
    for k in res.blocks:
        assert res.blk_qn_flow(k) == psi.blk_qn_flow(next(iter(psi.blocks)))
    assert two_site_matvec_flops(eff, psi) > 0


def test_rotated_contract_order():
    """Result axes: A's free axes cyclically after the run, then B's likewise."""
    rng = np.random.default_rng(1)
    from oracle.bsten import Index, IN, OUT
    ix = [Index.make([((0,), n)], IN) for n in (2, 3, 4, 5, 6)]
    A = BSTensor(ix)
    A.blocks[(0,) * 5] = rng.normal(size=(2, 3, 4, 5, 6))
    B = BSTensor([ix[4].inverse(), ix[0].inverse(), Index.make([((0,), 7)], OUT)])
    B.blocks[(0, 0, 0)] = rng.normal(size=(6, 2, 7))
    C = ContractRotated(A, B, 4, 0, 2)     # contract A axes (4, 0) with B axes (0, 1)
    ref = np.einsum("abcde,eag->bcdg", A.blocks[(0,) * 5], B.blocks[(0, 0, 0)])
    assert C.shape == (3, 4, 5, 7)
    assert np.allclose(C.blocks[(0, 0, 0, 0)], ref)


def test_env_update_hermitian_and_dense():
    m = heisenberg_chain(8)
    mpo = build_mpo(m)
    mps = random_mps(m, (0,), 12, seed=5)
    lenvs, renvs = init_envs(mps, mpo)
    L1 = UpdateSiteLenvs(lenvs[0], mps[0], mpo[0])
    Ld = np.einsum("xwa,xtb,wsty,asc->byc", lenvs[0].to_dense(), mps[0].to_dense().conj(), mpo[0].to_dense(),
                   mps[0].to_dense(), optimize=True)
    assert np.allclose(L1.to_dense(), Ld)
    R1 = renvs[1]
    Rd = np.einsum("bwc,xtb,ysta,zsc->xyz", renvs[0].to_dense(), mps[-1].to_dense().conj(), mpo[-1].to_dense(),
                   mps[-1].to_dense(), optimize=True)
    assert np.allclose(R1.to_dense(), Rd)
    # H_eff from true environments is Hermitian: <x|Hy> = <Hx|y>
    n = 8
    eff = [lenvs[0], mpo[0], mpo[1], renvs[n - 2]]
    psi = Contract(mps[0], mps[1], ([2], [0]))
    rng = np.random.default_rng(0)
    x, y = psi.copy(), psi.copy()
    for t in (x, y):
        for k in t.blocks:
            t.blocks[k] = rng.normal(size=t.blocks[k].shape)
    lhs = vdot(eff_ham_mul_two_site_state(eff, y), x)
    rhs = vdot(y, eff_ham_mul_two_site_state(eff, x))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_linear_combine_union_and_dag():
    from oracle.bsten import Index, IN, OUT
    i0 = Index.make([((0,), 2), ((1,), 3)], IN)
    A = BSTensor([i0, i0.inverse()])
    A.blocks[(0, 0)] = np.ones((2, 2))
    B = BSTensor([i0, i0.inverse()])
    B.blocks[(1, 1)] = 2 * np.ones((3, 3))
    res = A.copy()
    LinearCombine([3.0], [B], 0.5, res)
    assert set(res.blocks) == {(0, 0), (1, 1)}
    assert np.allclose(res.blocks[(0, 0)], 0.5) and np.allclose(res.blocks[(1, 1)], 6.0)
    d = Dag(A)
    assert [ix.dir for ix in d.indexes] == [OUT, IN]
