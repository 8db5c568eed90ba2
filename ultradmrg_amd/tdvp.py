"""GPU-resident two-site finite TDVP (real-time evolution of an MPS).

Follows ``include/qlmps/algorithm/tdvp/two_site_update_finite_tdvp_impl.h``:

* ``TwoSiteTDVP.two_site_evolve``  ``TwoSiteFiniteTDVPEvolution`` (:272-403):
  psi = A_l . A_r evolved by exp(-i step H_eff) through ``qlg_lanczos_expmv``,
  block SVD with the left site's divergence, U / S.V^H (right move) or
  U.S / V^H (left move), then the environment next to the moved centre.
* ``TwoSiteTDVP.one_site_backward`` ``SingleSiteFiniteTDVPBackwardEvolution``
  with dir 'n' (:405-531): the centre tensor evolves by exp(+i step H_1site).
* ``TwoSiteTDVP.sweep``            ``TwoSiteFiniteTDVPSweep`` (:207-271): right
  sweep i = 0..N-3 (two-site tau/2, backward at i+1), left sweep i = N-1..2
  (two-site tau/2, a full tau at i = N-1; backward at i-1), then two-site tau/2
  at (0, 1).  One call advances the state by tau.
* ``apply_local_op``               the operator insertion at ``site_0`` and the
  right-canonicalisation that ``TwoSiteFiniteTDVP`` performs (:123-160).
* ``local_expectations``           <O_i> for every site, from left transfer
  matrices (environment updates with a bond-dimension-1 MPO) with the centre
  moved to site 0.

Data is complex128; the Krylov exponential, SVD and environments run on the GPU.
The sector GEMMs of the Krylov exponentials can run in the split-BF16 modes
(``TDVPParams.precision``); the Krylov recurrence, the small-matrix
exponential, the SVD and the environments stay FP64.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .solver import EffectiveHamiltonian, LanczosParams
from .sweep import block_svd, center_to_zero, divergence, trivial_env
from .tensor import OUT, Axis, BlockTensor, contract, env_update


@dataclass
class TDVPParams:
    """The evolution part of ``algorithm/tdvp/tdvp_evolve_params.h:21-75``."""
    tau: float
    steps: int
    Dmin: int
    Dmax: int
    trunc_err: float
    lancz: LanczosParams
    verbose: bool = False
    precision: str = "fp64"   # sector GEMMs of the Krylov exponentials: "fp64", "bf16x6" or "bf16x3"


@dataclass
class TDVPStats:
    step_seconds: list = field(default_factory=list)
    expmv_s: float = 0.0
    svd_s: float = 0.0
    env_s: float = 0.0
    matvecs: int = 0


def to_complex(t: BlockTensor) -> BlockTensor:
    if t.data.is_complex():
        return t
    return BlockTensor(t.axes, t.coords, t.data.to(torch.complex128), t.offsets)


class TwoSiteTDVP:
    """Two-site TDVP on a GPU-resident MPS whose sites 1..N-1 are right-canonical."""

    def __init__(self, mps: List[BlockTensor], mpo: List[BlockTensor], params: TDVPParams):
        self.mps = [to_complex(t) for t in mps]
        self.mpo = [to_complex(t) for t in mpo]
        self.p = params
        self.stats = TDVPStats()
        self.rebuild_envs()

    def rebuild_envs(self):
        n, mps, mpo = len(self.mps), self.mps, self.mpo
        dev, dt = mps[0].data.device, mps[0].data.dtype
        self.renvs: List[Optional[BlockTensor]] = [None] * n
        self.lenvs: List[Optional[BlockTensor]] = [None] * n
        self.renvs[0] = trivial_env(mps[-1].axes[2], mpo[-1].axes[3], dev, dt)
        for i in range(1, n - 1):
            self.renvs[i] = env_update(self.renvs[i - 1], mps[n - i], mpo[n - i], "r")
        self.lenvs[0] = trivial_env(mps[0].axes[0], mpo[0].axes[0], dev, dt)

    def two_site_evolve(self, step: float, direction: str, target: int):
        n, p, st = len(self.mps), self.p, self.stats
        if direction == "r":
            l, r, llen, rlen = target, target + 1, target, n - (target + 2)
        else:
            l, r, llen, rlen = target - 1, target, target - 1, n - target - 1
        t0 = time.perf_counter()
        psi = contract(self.mps[l], self.mps[r], 1)
        H = EffectiveHamiltonian([self.lenvs[llen], self.mpo[l], self.mpo[r], self.renvs[rlen]], psi,
                                 precision=p.precision)
        res = H.expmv(psi, step, p.lancz)
        H.close()
        t1 = time.perf_counter()
        sv = block_svd(res.expmv, 2, divergence(self.mps[l]), p.trunc_err, p.Dmin, p.Dmax)
        t2 = time.perf_counter()
        if direction == "r":
            self.mps[l] = sv.u
            self.mps[r] = contract(sv.s, sv.vt, 1)
            self.lenvs[llen + 1] = env_update(self.lenvs[llen], self.mps[target], self.mpo[target], "l")
        else:
            self.mps[l] = contract(sv.u, sv.s, 1)
            self.mps[r] = sv.vt
            self.renvs[rlen + 1] = env_update(self.renvs[rlen], self.mps[target], self.mpo[target], "r")
        t3 = time.perf_counter()
        st.expmv_s += t1 - t0
        st.svd_s += t2 - t1
        st.env_s += t3 - t2
        st.matvecs += res.iters
        return sv

    def one_site_backward(self, step: float, target: int):
        n, p = len(self.mps), self.p
        t0 = time.perf_counter()
        H = EffectiveHamiltonian([self.lenvs[target], self.mpo[target], self.renvs[n - target - 1]],
                                 self.mps[target], precision=p.precision)
        res = H.expmv(self.mps[target], -step, p.lancz)
        H.close()
        self.mps[target] = res.expmv
        self.stats.expmv_s += time.perf_counter() - t0
        self.stats.matvecs += res.iters

    def sweep(self):
        """Advance the state by tau (``TwoSiteFiniteTDVPSweep``)."""
        n, tau = len(self.mps), self.p.tau
        t0 = time.perf_counter()
        for i in range(n - 2):
            self.two_site_evolve(tau / 2, "r", i)
            self.one_site_backward(tau / 2, i + 1)
        for i in range(n - 1, 1, -1):
            self.two_site_evolve(tau if i == n - 1 else tau / 2, "l", i)
            self.one_site_backward(tau / 2, i - 1)
        self.two_site_evolve(tau / 2, "r", 0)
        torch.cuda.synchronize()
        self.stats.step_seconds.append(time.perf_counter() - t0)

    def measure(self, op: np.ndarray) -> np.ndarray:
        """<O_i> for all sites, on a copy of the state with its centre moved to
        site 0 (the sweep leaves it at site 1); the evolving state and its
        environments are left as they are."""
        mps = list(self.mps)
        center_to_zero(mps)
        return local_expectations(mps, op)


def apply_local_op(mps: List[BlockTensor], op: np.ndarray, site: int):
    """mps[site] <- op on the physical index (op[s_in, s_out]; a charge-shifting
    operator is allowed: the site tensor's divergence shifts by its charge),
    then sites site .. 1 right-canonical (the charge moves to site 0)."""
    t = mps[site]
    phys = t.axes[1]
    off = np.concatenate([[0], np.cumsum(phys.degs)])
    dev, dt = t.data.device, t.data.dtype
    ob = {}
    for si in range(phys.nsect):
        for so in range(phys.nsect):
            blk = op[off[si]:off[si + 1], off[so]:off[so + 1]]
            if np.any(blk != 0):
                ob[(si, so)] = torch.from_numpy(np.ascontiguousarray(blk)).to(device=dev, dtype=dt)
    pieces = {}
    for i, c in enumerate(t.coords):
        blk = t.data[t.offsets[i]:t.offsets[i] + t.block_size(c)].view(t.block_shape(c))
        for (si, so), o in ob.items():
            if si != c[1]:
                continue
            piece = torch.einsum("asb,st->atb", blk, o)
            key = (c[0], so, c[2])
            pieces[key] = pieces[key] + piece if key in pieces else piece
    if not pieces:
        raise ValueError("the operator annihilates the state at this site")
    keys = sorted(pieces)
    mps[site] = BlockTensor(t.axes, keys, torch.cat([pieces[k].reshape(-1) for k in keys]))
    center_to_zero(mps, site)


def _site_mpo(op: np.ndarray, phys: Axis, dev, dt) -> BlockTensor:
    zero = tuple(0 for _ in phys.qns[0])
    w = Axis.make([(zero, 1)], OUT)
    off = np.concatenate([[0], np.cumsum(phys.degs)])
    blocks = {}
    for si in range(phys.nsect):
        for so in range(phys.nsect):
            blk = op[off[si]:off[si + 1], off[so]:off[so + 1]]
            if np.any(blk != 0):
                blocks[(0, si, so, 0)] = blk.reshape(1, phys.degs[si], phys.degs[so], 1)
    return BlockTensor.from_blocks([w.inverse(), phys.inverse(), phys, w], blocks, device=dev, dtype=dt)


def local_expectations(mps: List[BlockTensor], op: np.ndarray) -> np.ndarray:
    """<psi|O_i|psi> / <psi|psi> for every site of an MPS whose sites 1..N-1 are right-canonical."""
    phys = mps[0].axes[1]
    dev, dt = mps[0].data.device, mps[0].data.dtype
    Wo, Wi = _site_mpo(op, phys, dev, dt), _site_mpo(np.eye(op.shape[0]), phys, dev, dt)
    E = trivial_env(mps[0].axes[0], Wi.axes[0], dev, dt)
    out = []
    for t in mps:
        M = env_update(E, t, Wo, "l")
        acc = 0.0
        for i, c in enumerate(M.coords):
            if c[0] == c[2]:
                blk = M.data[M.offsets[i]:M.offsets[i] + M.block_size(c)].view(M.block_shape(c))
                acc += complex(torch.diagonal(blk[:, 0, :]).sum())
        out.append(acc)
        E = env_update(E, t, Wi, "l")
    nrm = 0.0   # <psi|psi>: truncation in the sweep does not renormalise the state
    for i, c in enumerate(E.coords):
        if c[0] == c[2]:
            blk = E.data[E.offsets[i]:E.offsets[i] + E.block_size(c)].view(E.block_shape(c))
            nrm += complex(torch.diagonal(blk[:, 0, :]).sum())
    return np.real(np.array(out)) / np.real(nrm)
