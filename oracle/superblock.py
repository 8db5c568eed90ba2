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

Parity status: pinned by the identity with the MPO-form matvec (itself pinned
by the reference's Lanczos KAT and energies) on identical operators.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .bsten import BSTensor, ContractRotated, Index

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
