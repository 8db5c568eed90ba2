"""Oracle: the reference's DMRG-form (super-block) effective Hamiltonian.

TEST INFRASTRUCTURE ONLY (see ``oracle/bsten.py``).

The DMRG executor (``include/qlmps/algorithm/dmrg/dmrg_impl.h:319-341``) groups
the two-site Hamiltonian by the MPO bond j between the two sites:

    H psi = sum_j [ sum_i L_i (x) W_l(i, j) ] psi [ sum_k W_r(j, k) (x) R_k ]

with rank-2 block operators L_i[a, a'], R_k[b, b'] (one per MPO bond state of
the left / right block) and rank-2 site operators W(i, j)[s, s'] (the non-null
entries of the MatReprMPO).  ``super_block_hamiltonian_mul_two_site_state``
restates ``include/qlmps/algorithm/dmrg/lanczos_dmrg_solver_impl.h:199-283``
term by term: per j, psi x L_i over a, x W_l(i, j) over s1, summed over i;
then x W_r(j, k) over s2, x R_k over b, summed over k; then summed over j.

``mpo_form_to_terms`` writes MPO-form operands (L[a', w, a], W_l[w, s, s', y],
W_r[y, s, s', z], R[b', z, b]) as such term groups, one rank-2 operator per
bond *element*, so the two forms can be checked against each other.

The DMRG executor around it is restated too: ``SparMat`` / ``mat_repr_mpo``
(the MatReprMPO, ``one_dim_tn/mpo/mat_repr_mpo.h:21`` with
``mpogen/symb_alg/sparse_mat.h:20-80``), ``UpdateLeftBlockOps`` /
``UpdateRightBlockOps`` (``algorithm/dmrg/dmrg_init.h:131-205``), the boundary
operator groups (``dmrg_init.h:207-309``), ``set_effective_hamiltonian_terms``
(``dmrg_impl.h:319-341``) and ``finite_dmrg`` (``DMRGSweep_`` /
``TwoSiteUpdate_``, ``dmrg_impl.h:172-317``, with left_boundary 0 and
right_boundary N-1).

Parity status: the matvec is pinned by the identity with the MPO-form matvec
(itself pinned by the reference's Lanczos KAT and energies) on identical
operators; the block-operator updates by the identity with the MPO-form
environment updates sliced per bond element; ``finite_dmrg`` by the
reference's FiniteDMRG known answers (test_finite_dmrg.cc) and exact
diagonalisation.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .bsten import BSTensor, Contract, ContractRotated, Dag, Index, inverse_index

Terms = List[Tuple[List[Tuple[BSTensor, BSTensor]], List[Tuple[BSTensor, BSTensor]]]]


def _add(acc, t):
    if acc is None:
        return t
    out = BSTensor(acc.indexes, np.result_type(acc.dtype, t.dtype))
    for k in set(acc.blocks) | set(t.blocks):
        a, b = acc.blocks.get(k), t.blocks.get(k)
        out.blocks[k] = a + b if (a is not None and b is not None) else (a if a is not None else b).copy()
    return out


def super_block_hamiltonian_mul_two_site_state(terms: Terms, state: BSTensor) -> BSTensor:
    """``lanczos_dmrg_solver_impl.h:199-283``: rotated contractions
    ``Contract<false, true>(x, op, 0, 0, 1)`` contract axis 0 of x with axis 0 of
    op and move op's free axis to the end."""
    res = None
    for bs_terms, sb_terms in terms:
        half = None
        for L_i, W_ij in bs_terms:                          # :200-217
            t = ContractRotated(state, L_i, 0, 0, 1)        # [s1, s2, b, a']
            half = _add(half, ContractRotated(t, W_ij, 0, 0, 1))   # [s2, b, a', s1']
        if half is None:
            continue
        full = None
        for W_jk, R_k in sb_terms:                          # :222-243
            t = ContractRotated(half, W_jk, 0, 0, 1)        # [b, a', s1', s2']
            full = _add(full, ContractRotated(t, R_k, 0, 0, 1))    # [a', s1', s2', b']
        if full is not None:
            res = _add(res, full)
    return res


def _bond_elements(ix: Index):
    out = []
    for s in range(ix.nsect):
        for p in range(ix.deg(s)):
            out.append((s, p))
    return out


def mpo_form_to_terms(L: BSTensor, Wl: BSTensor, Wr: BSTensor, R: BSTensor) -> Terms:
    """Split MPO-form operands into the reference's super-block term groups."""
    a_bra, w_ix, a_ket = L.indexes
    b_bra, z_ix, b_ket = R.indexes
    # L_i[a (OUT), a' (IN)] = L[a', i, a]^T
    Ls = {}
    for (s, p) in _bond_elements(w_ix):
        t = BSTensor([a_ket, a_bra], L.dtype)
        for (x, ws, z), blk in L.blocks.items():
            if ws == s:
                t.blocks[(z, x)] = blk[:, p, :].T.copy()
        Ls[(s, p)] = t
    Rs = {}
    for (s, p) in _bond_elements(z_ix):
        t = BSTensor([b_ket, b_bra], R.dtype)
        for (x, zs, z), blk in R.blocks.items():
            if zs == s:
                t.blocks[(z, x)] = blk[:, p, :].T.copy()
        Rs[(s, p)] = t

    def site_op(W, i, j):
        # W[wl, s_in, s_out, wr] element (i, j) -> rank-2 [s_in (IN), s_out (OUT)]
        t = BSTensor([W.indexes[1], W.indexes[2]], W.dtype)
        for (ws, si, so, wr), blk in W.blocks.items():
            if ws == i[0] and wr == j[0]:
                v = blk[i[1], :, :, j[1]]
                if np.any(v != 0):
                    t.blocks[(si, so)] = v.copy()
        return t if t.blocks else None

    y_ix = Wl.indexes[3]
    terms: Terms = []
    for j in _bond_elements(y_ix):
        bs = []
        for i in _bond_elements(w_ix):
            op = site_op(Wl, i, j)
            if op is not None and Ls[i].blocks:
                bs.append((Ls[i], op))
        sb = []
        for k in _bond_elements(z_ix):
            op = site_op(Wr, j, k)
            if op is not None and Rs[k].blocks:
                sb.append((op, Rs[k]))
        if bs and sb:
            terms.append((bs, sb))
    return terms


# --------------------------------------------------------------------------
# the DMRG executor (dmrg_impl.h, dmrg_init.h)
# --------------------------------------------------------------------------
@dataclass
class SparMat:
    """``sparse_mat.h:20-80``: a rows x cols matrix of rank-2 site operators
    [s_in (IN), s_out (OUT)]; missing entries are null."""
    rows: int
    cols: int
    data: Dict[Tuple[int, int], BSTensor] = field(default_factory=dict)

    def IsNull(self, i, j):
        return (i, j) not in self.data

    def __call__(self, i, j):
        return self.data[(i, j)]


def mat_repr_mpo(mpo: List[BSTensor]) -> List[SparMat]:
    """Each MPO tensor W[w_l, s_in, s_out, w_r] as its matrix of site
    operators over the bond elements (sector-major order), entries with a
    nonzero value only."""
    out = []
    for W in mpo:
        le, re = _bond_elements(W.indexes[0]), _bond_elements(W.indexes[3])
        sm = SparMat(len(le), len(re))
        for (ws, si, so, wr), blk in W.blocks.items():
            for a in range(blk.shape[0]):
                for z in range(blk.shape[3]):
                    v = blk[a, :, :, z]
                    if not np.any(v != 0):
                        continue
                    i, j = le.index((ws, a)), re.index((wr, z))
                    if (i, j) not in sm.data:
                        sm.data[(i, j)] = BSTensor([W.indexes[1], W.indexes[2]], W.dtype)
                    sm.data[(i, j)].blocks[(si, so)] = v.copy()
        out.append(sm)
    return out


def _id_scalar(ix: Index, dtype, left: bool) -> BSTensor:
    """``dmrg_init.h:220-225, 247-252``: the 1 x 1 identity {Inverse(ix), ix}
    (left) or {Inverse(ix), ix} with ix the last MPS tensor's right bond (right)."""
    t = BSTensor([inverse_index(ix), ix], dtype)
    t.blocks[(0, 0)] = np.ones((1, 1), dtype=dtype)
    return t


def _sum(acc, t):
    return t if acc is None else _add(acc, t)


def UpdateRightBlockOps(rog: List[BSTensor], mps: BSTensor, sm: SparMat) -> List[Optional[BSTensor]]:
    """``dmrg_init.h:131-161``: per row, the site-block operator
    sum_col W(row, col) (x) R_col, then contracted with the MPS tensor and its
    conjugate.  Rows with no non-null entry give None."""
    mps_dag = Dag(mps)
    out = []
    for row in range(sm.rows):
        sbo = None
        for col in range(sm.cols):
            if not sm.IsNull(row, col) and rog[col] is not None:
                sbo = _sum(sbo, Contract(sm(row, col), rog[col], ([], [])))
        if sbo is None:
            out.append(None)
            continue
        t = Contract(mps, sbo, ([1, 2], [0, 2]))
        out.append(Contract(t, mps_dag, ([1, 2], [1, 2])))
    return out


def UpdateLeftBlockOps(log: List[BSTensor], mps: BSTensor, sm: SparMat) -> List[Optional[BSTensor]]:
    """``dmrg_init.h:179-205``: per column, the block-site operator
    sum_row L_row (x) W(row, col), then contracted with the MPS tensor and its
    conjugate."""
    mps_dag = Dag(mps)
    out = []
    for col in range(sm.cols):
        bso = None
        for row in range(sm.rows):
            if not sm.IsNull(row, col) and log[row] is not None:
                bso = _sum(bso, Contract(log[row], sm(row, col), ([], [])))
        if bso is None:
            out.append(None)
            continue
        t = Contract(bso, mps, ([0, 2], [0, 1]))
        out.append(Contract(t, mps_dag, ([0, 1], [0, 1])))
    return out


def set_effective_hamiltonian_terms(lops, rops, sm_l: SparMat, sm_r: SparMat) -> Terms:
    """``dmrg_impl.h:319-341`` (terms with a None block operator, which
    contribute nothing, are left out)."""
    terms: Terms = []
    for j in range(sm_l.cols):
        bs = [(lops[i], sm_l(i, j)) for i in range(len(lops))
              if not sm_l.IsNull(i, j) and lops[i] is not None]
        sb = [(sm_r(j, k), rops[k]) for k in range(sm_r.cols)
              if not sm_r.IsNull(j, k) and rops[k] is not None]
        terms.append((bs, sb))
    return terms


def finite_dmrg(mps: List[BSTensor], mro: List[SparMat], p, log=None) -> float:
    """``FiniteDMRG`` / ``DMRGExecutor::Execute`` (``dmrg_impl.h:140-202``) on a
    right-canonical MPS: right block-operator groups for every length
    (``InitBlockOps``, ``dmrg_init.h:280-309``), then ``p.sweeps`` sweeps of
    ``TwoSiteUpdate_`` (``:213-317``) with left_boundary 0, right_boundary N-1."""
    from .dmrg import block_svd
    from .lanczos import LanczosSolver
    n = len(mps)
    dt = mps[0].dtype
    ropg: List[Optional[list]] = [None] * n
    lopg: List[Optional[list]] = [None] * n
    ropg[0] = [_id_scalar(mps[-1].indexes[2], dt, False)]
    for i in range(1, n - 1):
        ropg[i] = UpdateRightBlockOps(ropg[i - 1], mps[n - i], mro[n - i])
    lopg[0] = [_id_scalar(mps[0].indexes[0], dt, True)]
    e0 = None

    def update(direction, l, r):
        llen, rlen = l, n - 1 - r
        terms = set_effective_hamiltonian_terms(lopg[llen], ropg[rlen], mro[l], mro[r])
        psi = Contract(mps[l], mps[r], ([2], [0]))
        res = LanczosSolver(terms, psi, super_block_hamiltonian_mul_two_site_state, p.lancz)
        sv = block_svd(res.gs_vec, 2, mps[l].div(), p.trunc_err, p.Dmin, p.Dmax)
        if direction == "r":
            mps[l] = sv.u
            mps[r] = Contract(sv.s, sv.vt, ([1], [0]))
            lopg[llen + 1] = UpdateLeftBlockOps(lopg[llen], mps[l], mro[l])
        else:
            mps[l] = Contract(sv.u, sv.s, ([2], [0]))
            mps[r] = sv.vt
            ropg[rlen + 1] = UpdateRightBlockOps(ropg[rlen], mps[r], mro[r])
        if log is not None:
            log.append(dict(l=l, r=r, dir=direction, e0=res.gs_eng, iters=res.iters, D=sv.D))
        return res.gs_eng

    for _ in range(p.sweeps):
        for i in range(0, n - 2):
            e0 = update("r", i, i + 1)
        for i in range(n - 1, 1, -1):
            e0 = update("l", i - 1, i)
    return e0
