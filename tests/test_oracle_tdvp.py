"""Oracle two-site TDVP against exact evolution.

With bond dimensions at their maximum (every charge sector's Schmidt rank),
two-site TDVP integrates the Schroedinger equation exactly up to the Krylov
tolerance, so <O_i(t)> must match exp(-iHt)|psi0> of the dense Hamiltonian."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.bsten import BSTensor
from oracle.dmrg import init_envs, random_mps
from oracle.lanczos import LanczosParams
from oracle.models import build_mpo, dense_hamiltonian_sparse, heisenberg_chain, hubbard_cylinder
from oracle.tdvp import local_expectations, tdvp_sweep


def mps_to_vector(mps, qns, total):
    t = mps[0]
    acc = t.to_dense()                              # [1, d, D1]
    for s in mps[1:]:
        acc = np.tensordot(acc, s.to_dense(), axes=([-1], [0]))
    v = acc.reshape(-1)
    labels = np.array(qns)
    tot = np.zeros((1, labels.shape[1]), dtype=np.int64)
    for _ in range(len(mps)):
        tot = (tot[:, None, :] + labels[None, :, :]).reshape(-1, labels.shape[1])
    keep = np.where(np.all(tot == np.array(total)[None, :], axis=1))[0]
    return v, keep


def site_op_dense(op, site, n, d, keep):
    mats = [sp.identity(d, format="csr")] * n
    mats = list(mats)
    mats[site] = sp.csr_matrix(op.T)
    m = mats[0]
    for k in range(1, n):
        m = sp.kron(m, mats[k], format="csr")
    return m[keep][:, keep]


@pytest.mark.parametrize("which", ["heisenberg", "hubbard"])
def test_tdvp_exact_evolution(which):
    if which == "heisenberg":
        m, total, D, opname = heisenberg_chain(6), (0,), 8, "sz"
    else:
        m, total, D, opname = hubbard_cylinder(2, 2), (4, 0), 16, "nf"
    n, d = m.n, len(m.qns)
    mpo = build_mpo(m)
    mps = random_mps(m, total, D, seed=7)
    for i, t in enumerate(mps):                     # complex state
        c = BSTensor(t.indexes, np.complex128, t.div())
        c.blocks = {k: v.astype(np.complex128) for k, v in t.blocks.items()}
        mps[i] = c
    nrm = np.sqrt(sum(float(np.sum(np.abs(v) ** 2)) for v in mps[0].blocks.values()))
    mps[0].blocks = {k: v / nrm for k, v in mps[0].blocks.items()}
    v0, keep = mps_to_vector(mps, m.qns, total)
    psi0 = v0[keep]
    assert abs(np.linalg.norm(psi0) - 1) < 1e-12 and np.linalg.norm(np.delete(v0, keep)) < 1e-12
    H = dense_hamiltonian_sparse(m, total).astype(np.complex128)
    lenvs, renvs = init_envs(mps, mpo)
    tau, steps = 0.05, 4
    lp = LanczosParams(1e-12, 40)
    for _ in range(steps):
        tdvp_sweep(mps, lenvs, renvs, mpo, tau, 0.0, 1, 10 ** 6, lp)
    # the sweep ends with the centre at site 1: move it to site 0 for the measurement
    from oracle.tdvp import center_to_zero
    center_to_zero(mps)
    got = local_expectations(mps, m.ops[opname])
    psit = spla.expm_multiply(-1j * tau * steps * H, psi0)
    ref = [float(np.real(np.vdot(psit, site_op_dense(m.ops[opname], i, n, d, keep) @ psit))) for i in range(n)]
    assert np.abs(got - np.array(ref)).max() < 1e-9, (got, ref)
