"""Oracle: two-site finite DMRG (VMPS form) and two-site TDVP drivers, MPS
initialisation, environment updates and block SVD truncation.

TEST INFRASTRUCTURE ONLY (see ``oracle/bsten.py``).

Reference lines followed:

* ``UpdateSiteLenvs`` / ``UpdateSiteRenvs``  ``include/qlmps/algorithm/vmps/vmps_init.h:343-373``
* trivial boundary environments           ``vmps_init.h:157-187``
* ``TwoSiteFiniteVMPSSweep``               ``algorithm/vmps/two_site_update_finite_vmps_impl.h:197-222``
* ``TwoSiteFiniteVMPSUpdate``              ``two_site_update_finite_vmps_impl.h:232-373``
* ``DirectStateInitMps``                   ``one_dim_tn/mps/finite_mps/finite_mps_init.h:213-262``
* two-site TDVP sweep                      ``algorithm/tdvp/two_site_update_finite_tdvp_impl.h:204-531``

The block SVD truncation rule is this package's specification (qlten's rule is
not in the reference): singular values of all charge sectors are pooled and
sorted; D is the smallest count whose discarded weight sum(s^2)/total is
<= trunc_err, clamped to [Dmin, Dmax]; the reported truncation error is the
discarded weight at that D.  Parity of that rule is unpinned by the
reference; end-to-end energies pin the sweep.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Sequence

import numpy as np

from .bsten import IN, OUT, BSTensor, Contract, ContractRotated, Dag, Index, qn_add, qn_sub
from .lanczos import LanczosParams, LanczosSolver, eff_ham_mul_two_site_state
from .models import ModelSpec, phys_index


# --------------------------------------------------------------------------
# environments
# --------------------------------------------------------------------------
def UpdateSiteLenvs(lenv: BSTensor, mps_ten: BSTensor, mpo_ten: BSTensor) -> BSTensor:
    """``vmps_init.h:359-373``: L'[b*, w1, b] = A* L W A."""
    t = ContractRotated(lenv, mps_ten, 2, 0, 1)
    t2 = ContractRotated(t, mpo_ten, 1, 0, 2)
    return ContractRotated(Dag(mps_ten), t2, 0, 1, 2)


def UpdateSiteRenvs(renv: BSTensor, mps_ten: BSTensor, mpo_ten: BSTensor) -> BSTensor:
    """``vmps_init.h:343-357``: R'[a*, w0, a] = B* R W B."""
    t = ContractRotated(Dag(mps_ten), renv, 2, 0, 1)
    t2 = ContractRotated(t, mpo_ten, 1, 2, 2)
    return ContractRotated(t2, mps_ten, 3, 1, 2)


def trivial_env(mps_bond: Index, mpo_bond: Index, dtype) -> BSTensor:
    """``vmps_init.h:160-164``: T({mps_idx, Inverse(mpo_idx), Inverse(mps_idx)}), T(0,0,0)=1."""
    t = BSTensor([mps_bond, mpo_bond.inverse(), mps_bond.inverse()], dtype)
    t[(0, 0, 0)] = 1.0
    return t


# --------------------------------------------------------------------------
# block SVD with truncation
# --------------------------------------------------------------------------
@dataclass
class SVDRes:
    u: BSTensor
    s: BSTensor
    vt: BSTensor
    trunc_err: float
    D: int


def block_svd(t: BSTensor, ldims: int, ldiv, trunc_err: float, Dmin: int, Dmax: int) -> SVDRes:
    """Split ``t`` into left axes [0, ldims) and the rest; U carries divergence ``ldiv``."""
    left, right = t.indexes[:ldims], t.indexes[ldims:]
    ldiv = tuple(ldiv)

    def mid_qn(coords):
        # div(u) = sum_OUT - sum_IN over left axes + q_mid (mid OUT) = ldiv
        q = tuple(0 for _ in ldiv)
        for ix, s in zip(left, coords[:ldims]):
            q = qn_add(q, ix.qn(s)) if ix.dir == OUT else qn_sub(q, ix.qn(s))
        return qn_sub(ldiv, q)

    groups: Dict[tuple, dict] = {}
    for k in sorted(t.blocks):
        q = mid_qn(k)
        g = groups.setdefault(q, {"rows": [], "cols": []})
        lk, rk = k[:ldims], k[ldims:]
        if lk not in g["rows"]:
            g["rows"].append(lk)
        if rk not in g["cols"]:
            g["cols"].append(rk)
    dec = {}
    allsv = []
    for q, g in groups.items():
        rdims = [int(np.prod([ix.deg(s) for ix, s in zip(left, lk)])) for lk in g["rows"]]
        cdims = [int(np.prod([ix.deg(s) for ix, s in zip(right, rk)])) for rk in g["cols"]]
        M = np.zeros((sum(rdims), sum(cdims)), dtype=t.dtype)
        ro = np.concatenate([[0], np.cumsum(rdims)])
        co = np.concatenate([[0], np.cumsum(cdims)])
        for i, lk in enumerate(g["rows"]):
            for j, rk in enumerate(g["cols"]):
                b = t.blocks.get(lk + rk)
                if b is not None:
                    M[ro[i]:ro[i + 1], co[j]:co[j + 1]] = b.reshape(rdims[i], cdims[j])
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        dec[q] = (U, S, Vh, ro, co, rdims, cdims)
        allsv.extend((float(s), q, n) for n, s in enumerate(S))
    allsv.sort(key=lambda x: -x[0])
    sq = np.array([x[0] ** 2 for x in allsv])
    total = float(sq.sum())
    tail = np.concatenate([np.cumsum(sq[::-1])[::-1], [0.0]])  # tail[k] = sum_{i>=k}
    D = len(allsv)
    for k in range(len(allsv) + 1):
        if tail[k] / total <= trunc_err:
            D = k
            break
    D = max(D, min(Dmin, len(allsv)))
    D = min(D, Dmax, len(allsv))
    D = max(D, 1)
    err = float(tail[D] / total)
    keep: Dict[tuple, int] = {}
    for s, q, n in allsv[:D]:
        keep[q] = keep.get(q, 0) + 1
    qs = sorted(keep)
    mid_out = Index.make([(q, keep[q]) for q in qs], OUT)
    mid_in = mid_out.inverse()
    u = BSTensor(list(left) + [mid_out], t.dtype, ldiv)
    s = BSTensor([mid_in, mid_out], np.float64)
    vt = BSTensor([mid_in] + list(right), t.dtype)
    for mi, q in enumerate(qs):
        U, S, Vh, ro, co, rdims, cdims = dec[q]
        k = keep[q]
        g = groups[q]
        s.blocks[(mi, mi)] = np.diag(S[:k]).astype(np.float64)
        for i, lk in enumerate(g["rows"]):
            shp = tuple(ix.deg(x) for ix, x in zip(left, lk)) + (k,)
            u.blocks[lk + (mi,)] = U[ro[i]:ro[i + 1], :k].reshape(shp).copy()
        for j, rk in enumerate(g["cols"]):
            shp = (k,) + tuple(ix.deg(x) for ix, x in zip(right, rk))
            vt.blocks[(mi,) + rk] = Vh[:k, co[j]:co[j + 1]].reshape(shp).copy()
    return SVDRes(u, s, vt, err, D)


def entanglement_entropy(s: BSTensor) -> float:
    sv = np.concatenate([np.diag(b) for b in s.blocks.values()])
    p = sv ** 2
    p = p / p.sum()
    p = p[p > 0]
    return float(-(p * np.log(p)).sum())


# --------------------------------------------------------------------------
# MPS construction
# --------------------------------------------------------------------------
def direct_state_mps(m: ModelSpec, stat_labs: Sequence[int], dtype=np.float64) -> List[BSTensor]:
    """``finite_mps_init.h:213-262``: product state; site 0 carries the total charge."""
    n, qns = m.n, m.qns
    pout = phys_index(qns, OUT)
    zero = tuple(0 for _ in qns[0])
    div = zero
    for i in range(n):
        div = qn_add(div, qns[stat_labs[i]])
    mps = []
    lq = zero
    rq = qn_sub(div, qns[stat_labs[0]])
    lvb = Index.make([(zero, 1)], IN)
    for i in range(n):
        if i > 0:
            rq = qn_sub(lq, qns[stat_labs[i]]) if i < n - 1 else zero
        rvb = Index.make([(rq, 1)], OUT)
        t = BSTensor([lvb, pout, rvb], dtype)
        t[(0, stat_labs[i], 0)] = 1.0
        mps.append(t)
        lvb = rvb.inverse()
        lq = rq
    return mps


def _count_charges(qns, nsites):
    cnt = {tuple(0 for _ in qns[0]): 1}
    for _ in range(nsites):
        new = {}
        for q, c in cnt.items():
            for s in qns:
                k = qn_add(q, s)
                new[k] = new.get(k, 0) + c
        cnt = new
    return cnt


def random_mps(m: ModelSpec, total, D: int, seed: int, dtype=np.float64) -> List[BSTensor]:
    """Seeded random MPS in the sector ``total`` with bond dimension <= D,
    right-canonical from site 1 (centre at site 0).

    Bond b (right of site b) carries charge q with q = total - (charges of sites
    0..b) = charges of sites b+1..n-1; its degeneracy is min(#left, #right)
    scaled so that the bond holds at most D states.
    """
    n, qns = m.n, m.qns
    total = tuple(total)
    rng = np.random.default_rng(seed)
    pout = phys_index(qns, OUT)
    zero = tuple(0 for _ in qns[0])
    bonds = []
    for b in range(n - 1):
        cl = _count_charges(qns, b + 1)
        cr = _count_charges(qns, n - b - 1)
        degs = {}
        for x, c in cl.items():
            q = qn_sub(total, x)
            if q in cr:
                degs[q] = min(c, cr[q])
        tot = sum(degs.values())
        if tot > D:
            scale = D / tot
            degs = {q: max(1, int(round(v * scale))) for q, v in degs.items()}
            while sum(degs.values()) > D:
                qmax = max(degs, key=lambda q: degs[q])
                degs[qmax] -= 1
        bonds.append(Index.make(sorted(degs.items()), OUT))
    mps = []
    lvb = Index.make([(zero, 1)], IN)
    for i in range(n):
        rvb = bonds[i] if i < n - 1 else Index.make([(zero, 1)], OUT)
        t = BSTensor([lvb, pout, rvb], dtype, div=total if i == 0 else zero)
        for coords in t.allowed_blocks(total if i == 0 else zero):
            t.blocks[coords] = rng.uniform(-1, 1, size=t.blk_shape(coords)).astype(dtype)
        mps.append(t)
        lvb = rvb.inverse()
    right_canonicalize(mps, 1)
    nrm = mps[0].norm()
    mps[0].scale(1.0 / nrm)
    return mps


def right_canonicalize(mps: List[BSTensor], upto: int):
    """Make sites n-1 .. upto right-canonical (SVD without truncation)."""
    n = len(mps)
    for i in range(n - 1, upto - 1, -1):
        t = mps[i]
        ldiv = t.div() if t.div() is not None else tuple(0 for _ in t.indexes[0].qn(0))
        r = block_svd(t, 1, ldiv, 0.0, 1, 10 ** 9, )
        mps[i] = r.vt
        us = Contract(r.u, r.s, ([1], [0]))
        mps[i - 1] = Contract(mps[i - 1], us, ([2], [0]))


# --------------------------------------------------------------------------
# two-site DMRG (VMPS form)
# --------------------------------------------------------------------------
@dataclass
class SweepParams:
    """``algorithm/finite_vmps_sweep_params.h:31-49``."""
    sweeps: int
    Dmin: int
    Dmax: int
    trunc_err: float
    lancz: LanczosParams


def init_envs(mps, mpo):
    n = len(mps)
    dtype = mps[0].dtype
    renvs = [None] * n
    renvs[0] = trivial_env(mps[-1].indexes[2], mpo[-1].indexes[3], dtype)
    for i in range(1, n - 1):
        renvs[i] = UpdateSiteRenvs(renvs[i - 1], mps[n - i], mpo[n - i])
    lenvs = [None] * n
    lenvs[0] = trivial_env(mps[0].indexes[0], mpo[0].indexes[0], dtype)
    return lenvs, renvs


def two_site_update(mps, lenvs, renvs, mpo, p: SweepParams, direction: str, target: int, log=None,
                    solver=None):
    n = len(mps)
    if direction == "r":
        l, r = target, target + 1
        llen, rlen = target, n - (target + 2)
    else:
        l, r = target - 1, target
        llen, rlen = target - 1, n - target - 1
    eff = [lenvs[llen], mpo[l], mpo[r], renvs[rlen]]
    psi = Contract(mps[l], mps[r], ([2], [0]))
    if solver is None:
        res = LanczosSolver(eff, psi, eff_ham_mul_two_site_state, p.lancz)
    else:
        res = solver(eff, psi, p.lancz)
    ldiv = mps[l].div()
    sv = block_svd(res.gs_vec, 2, ldiv, p.trunc_err, p.Dmin, p.Dmax)
    if direction == "r":
        mps[l] = sv.u
        mps[r] = Contract(sv.s, sv.vt, ([1], [0]))
        lenvs[llen + 1] = UpdateSiteLenvs(lenvs[llen], mps[target], mpo[target])
    else:
        mps[l] = Contract(sv.u, sv.s, ([2], [0]))
        mps[r] = sv.vt
        renvs[rlen + 1] = UpdateSiteRenvs(renvs[rlen], mps[target], mpo[target])
    if log is not None:
        log.append(dict(site=target, dir=direction, e0=res.gs_eng, iters=res.iters, D=sv.D,
                        trunc_err=sv.trunc_err))
    return res.gs_eng


def two_site_vmps(mps, mpo, p: SweepParams, log=None, solver=None) -> float:
    """``TwoSiteFiniteVMPS`` (``two_site_update_finite_vmps_impl.h:95-125, 197-222``)."""
    n = len(mps)
    lenvs, renvs = init_envs(mps, mpo)
    e0 = None
    for _ in range(p.sweeps):
        for i in range(0, n - 2):
            e0 = two_site_update(mps, lenvs, renvs, mpo, p, "r", i, log, solver)
        for i in range(n - 1, 1, -1):
            e0 = two_site_update(mps, lenvs, renvs, mpo, p, "l", i, log, solver)
    return e0
