"""Oracle: two-site finite TDVP (real-time evolution of an MPS).

TEST INFRASTRUCTURE ONLY (see ``oracle/bsten.py``).

Restates ``include/qlmps/algorithm/tdvp/two_site_update_finite_tdvp_impl.h``:

* ``two_site_evolve``  ``TwoSiteFiniteTDVPEvolution`` (:272-403): psi = A_l A_r,
  psi <- exp(-i tau/2 H_eff) psi (``LanczosExpmvSolver``), SVD with the left
  site's divergence, U / S.V^H (right move) or U.S / V^H (left move), then the
  environment next to the moved centre.
* ``one_site_backward`` ``SingleSiteFiniteTDVPBackwardEvolution`` with dir 'n'
  (:405-531): the centre tensor evolves by exp(+i tau/2 H_1site) in place.
* ``tdvp_sweep``        ``TwoSiteFiniteTDVPSweep`` (:207-271): right sweep
  i = 0..N-3 (two-site tau/2, backward at i+1), left sweep i = N-1..2 (two-site
  tau/2, a full tau at i = N-1; backward at i-1), then two-site tau/2 at (0, 1).
* ``apply_local_op``    the operator insertion and right-canonicalisation
  from ``site_0`` down to 1 that ``TwoSiteFiniteTDVP`` performs (:123-160).
* ``local_expectations`` <psi|O_i|psi> / <psi|psi> for every site, from left
  transfer matrices with the centre moved to site 0.

Parity status: pinned against exact evolution exp(-iHt)|psi0> of the dense
Hamiltonian on small systems with untruncated bond dimension (two-site TDVP is
then exact up to the Krylov tolerance; tests/test_oracle_tdvp.py).  The
reference's own TDVP test measures a fermionic dynamic correlation
(tests/test_algorithm/test_two_site_finite_tdvp.cc:82-100) that needs
two-state fermionic overlaps, which this restatement does not cover.
"""
from __future__ import annotations

from typing import List

import numpy as np

from .bsten import BSTensor, Contract, Index
from .dmrg import UpdateSiteLenvs, UpdateSiteRenvs, block_svd, trivial_env
from .lanczos import (LanczosExpmvSolver, LanczosParams, eff_ham_mul_single_site_state,
                      eff_ham_mul_two_site_state)

OUT, IN = +1, -1


def _div(t: BSTensor):
    d = t.div()
    return d if d is not None else tuple(0 for _ in t.indexes[0].qn(0))


def two_site_evolve(mps, lenvs, renvs, mpo, step, trunc, Dmin, Dmax, lp: LanczosParams, direction, target):
    n = len(mps)
    if direction == "r":
        l, r, llen, rlen = target, target + 1, target, n - (target + 2)
    else:
        l, r, llen, rlen = target - 1, target, target - 1, n - target - 1
    psi = Contract(mps[l], mps[r], ([2], [0]))
    res = LanczosExpmvSolver([lenvs[llen], mpo[l], mpo[r], renvs[rlen]], psi, eff_ham_mul_two_site_state, step, lp)
    sv = block_svd(res.expmv, 2, _div(mps[l]), trunc, Dmin, Dmax)
    if direction == "r":
        mps[l] = sv.u
        mps[r] = Contract(sv.s, sv.vt, ([1], [0]))
        lenvs[llen + 1] = UpdateSiteLenvs(lenvs[llen], mps[target], mpo[target])
    else:
        mps[l] = Contract(sv.u, sv.s, ([2], [0]))
        mps[r] = sv.vt
        renvs[rlen + 1] = UpdateSiteRenvs(renvs[rlen], mps[target], mpo[target])
    return res.iters


def one_site_backward(mps, lenvs, renvs, mpo, step, lp: LanczosParams, target):
    n = len(mps)
    res = LanczosExpmvSolver([lenvs[target], mpo[target], renvs[n - target - 1]], mps[target],
                             eff_ham_mul_single_site_state, -step, lp)
    mps[target] = res.expmv
    return res.iters


def tdvp_sweep(mps, lenvs, renvs, mpo, tau, trunc, Dmin, Dmax, lp: LanczosParams):
    n = len(mps)
    for i in range(n - 2):
        two_site_evolve(mps, lenvs, renvs, mpo, tau / 2, trunc, Dmin, Dmax, lp, "r", i)
        one_site_backward(mps, lenvs, renvs, mpo, tau / 2, lp, i + 1)
    for i in range(n - 1, 1, -1):
        two_site_evolve(mps, lenvs, renvs, mpo, tau if i == n - 1 else tau / 2, trunc, Dmin, Dmax, lp, "l", i)
        one_site_backward(mps, lenvs, renvs, mpo, tau / 2, lp, i - 1)
    two_site_evolve(mps, lenvs, renvs, mpo, tau / 2, trunc, Dmin, Dmax, lp, "r", 0)


def apply_local_op(mps: List[BSTensor], op: BSTensor, site: int):
    """mps[site] <- op acting on the physical index (op[s_in (IN), s_out (OUT)]),
    then sites site .. 1 right-canonical (the centre moves to site 0)."""
    t = Contract(mps[site], op, ([1], [0]))           # [a, b, s']
    mps[site] = t.transpose([0, 2, 1])
    for i in range(site, 0, -1):
        r = block_svd(mps[i], 1, _div(mps[i]), 0.0, 1, 10 ** 9)
        mps[i] = r.vt
        mps[i - 1] = Contract(mps[i - 1], Contract(r.u, r.s, ([1], [0])), ([2], [0]))


def _site_mpo(op: np.ndarray, phys: Index, dtype) -> BSTensor:
    """A bond-dimension-1 MPO tensor W[wl, s_in, s_out, wr] for a charge-neutral site operator."""
    zero = tuple(0 for _ in phys.qn(0))
    w = Index.make([(zero, 1)], OUT)
    t = BSTensor([w.inverse(), phys.inverse(), phys, w], dtype)
    offs = phys.offsets()
    for si in range(phys.nsect):
        for so in range(phys.nsect):
            blk = op[offs[si]:offs[si] + phys.deg(si), offs[so]:offs[so] + phys.deg(so)]
            if np.any(blk != 0):
                t.blocks[(0, si, so, 0)] = blk.reshape(1, phys.deg(si), phys.deg(so), 1).astype(dtype)
    return t


def local_expectations(mps: List[BSTensor], op: np.ndarray) -> np.ndarray:
    """<psi|O_i|psi> / <psi|psi> for every site i of an MPS whose sites
    1..N-1 are right-canonical (centre at site 0)."""
    phys = mps[0].indexes[1]
    dt = np.result_type(mps[0].dtype, np.float64)
    Wo, Wi = _site_mpo(op, phys, dt), _site_mpo(np.eye(op.shape[0]), phys, dt)
    E = trivial_env(mps[0].indexes[0], Wi.indexes[0], dt)
    out = []
    for t in mps:
        M = UpdateSiteLenvs(E, t, Wo)
        out.append(sum(np.trace(v[:, 0, :]) for k, v in M.blocks.items() if k[0] == k[2]))
        E = UpdateSiteLenvs(E, t, Wi)
    nrm = sum(np.trace(v[:, 0, :]) for k, v in E.blocks.items() if k[0] == k[2])
    return np.real(np.array(out)) / np.real(nrm)


def center_to_zero(mps: List[BSTensor]):
    """Right-canonicalise sites N-1 .. 1 (SVD without truncation)."""
    n = len(mps)
    for i in range(n - 1, 0, -1):
        r = block_svd(mps[i], 1, _div(mps[i]), 0.0, 1, 10 ** 9)
        mps[i] = r.vt
        mps[i - 1] = Contract(mps[i - 1], Contract(r.u, r.s, ([1], [0])), ([2], [0]))
