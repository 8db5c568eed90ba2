"""MPO construction of the package (``ultradmrg_amd.mpo.build_mpo``), on the CPU.

The reference builds MPOs with its MPO generator
(one_dim_tn/mpo/mpogen/mpogen_impl.h:443-504); the package builds its own
block-sparse MPO from the same term list.  These tests contract the package's
MPO, built on the CPU, into a dense matrix and compare it with the oracle's
Hamiltonian summed term by term as Kronecker products
(``oracle.models.dense_hamiltonian_sparse``), which uses no MPO at all."""
import numpy as np
import pytest
import torch

from oracle import models as O
from ultradmrg_amd import mpo as M


def _dense(t):
    starts = [np.concatenate([[0], np.cumsum(ax.degs)[:-1]]).astype(int) for ax in t.axes]
    out = np.zeros([ax.dim for ax in t.axes], dtype=complex if t.is_complex else float)
    for c, blk in t.blocks().items():
        out[tuple(slice(o[s], o[s] + ax.degs[s]) for o, ax, s in zip(starts, t.axes, c))] = blk
    return out


def _mpo_matrix(mpo):
    """<out|H|in> from site tensors W[left, in, out, right] (site 0 most significant)."""
    T = None
    for W in mpo:
        w = _dense(W)
        if T is None:
            assert w.shape[0] == 1
            T = w[0]
            continue
        d = w.shape[1]
        T = np.einsum("ioa,ajkb->ijokb", T, w).reshape(T.shape[0] * d, T.shape[1] * d, w.shape[3])
    assert T.shape[2] == 1
    return T[:, :, 0].T


@pytest.mark.parametrize("n", [2, 3, 5])
def test_heisenberg_mpo_matches_term_sum(n):
    fields = list(np.linspace(-0.7, 0.9, n))
    H = _mpo_matrix(M.build_mpo(M.heisenberg_chain(n, J=1.3, fields=fields), device="cpu"))
    ref = O.dense_hamiltonian_sparse(O.heisenberg_chain(n, J=1.3, fields=fields)).toarray()
    np.testing.assert_allclose(H, ref, atol=1e-13)


@pytest.mark.parametrize("lx,ly,t2", [(2, 2, 0.0), (2, 2, 0.3), (3, 1, 0.0)])
def test_hubbard_mpo_matches_term_sum(lx, ly, t2):
    H = _mpo_matrix(M.build_mpo(M.hubbard_cylinder(lx, ly, t=1.0, U=4.0, t2=t2), device="cpu"))
    ref = O.dense_hamiltonian_sparse(O.hubbard_cylinder(lx, ly, t=1.0, U=4.0, t2=t2)).toarray()
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(H, ref, atol=1e-13)


@pytest.mark.parametrize("n", [2, 4])
def test_complex_coefficients_keep_operator_orientation(n):
    """A Hermitian H with complex couplings is not symmetric (H^T = conj(H)), so
    a transposed operator convention or a conjugated coefficient changes it."""
    m = M.heisenberg_chain(n, J=0.8)
    spec = O.heisenberg_chain(n, J=0.8)
    for i in range(n - 1):
        c = 0.4j * (i + 1) + 0.1
        m.add(c, (i, "sp"), (i + 1, "sm"))
        m.add(np.conj(c), (i, "sm"), (i + 1, "sp"))
        spec.terms.append(O.Term(c, [(i, "sp"), (i + 1, "sm")]))
        spec.terms.append(O.Term(np.conj(c), [(i, "sm"), (i + 1, "sp")]))
    H = _mpo_matrix(M.build_mpo(m, device="cpu", dtype=torch.complex128))
    ref = O.dense_hamiltonian_sparse(spec).toarray()
    assert np.abs(ref - ref.T).max() > 0.1 and np.allclose(ref, ref.conj().T)
    np.testing.assert_allclose(H, ref, atol=1e-13)
