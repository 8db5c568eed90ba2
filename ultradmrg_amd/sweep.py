"""GPU-resident two-site finite DMRG (the reference's VMPS form): MPS
initialisation, environments, block SVD truncation and the sweep driver.

Reference behaviour followed:
* ``TwoSiteFiniteVMPS`` / ``TwoSiteFiniteVMPSSweep`` / ``TwoSiteFiniteVMPSUpdate``
  (``include/qlmps/algorithm/vmps/two_site_update_finite_vmps_impl.h:95-373``):
  right-moving updates i = 0..N-3, left-moving i = N-1..2; psi0 = A_l . A_r;
  Lanczos; SVD with the left divergence of A_l; the environment next to the
  updated pair is refreshed; one log line per update.
* trivial boundary environments (``vmps_init.h:157-187``).
* ``DirectStateInitMps`` (``one_dim_tn/mps/finite_mps/finite_mps_init.h:213-262``).

Everything stays on the GPU: contractions and environment updates go through
the library's planners, the Krylov solve through ``qlg_lanczos_gs``, and the
per-sector SVDs through batched FP64 eigen-decompositions of the Gram matrix
on each sector matrix's smaller side (torch / rocSOLVER).  The truncation
rule is this build's: pooled singular values, smallest D whose discarded
weight sum(s^2)/total <= trunc_err, clamped to [Dmin, Dmax].
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .mpo import Model, physical_axis
from .solver import EffectiveHamiltonian, LanczosParams
from .tensor import IN, OUT, Axis, BlockTensor, _flow, contract, env_update


def divergence(t: BlockTensor) -> Tuple[int, ...]:
    return _flow(t.axes, t.coords[0])


def block_view(t: BlockTensor, i: int) -> torch.Tensor:
    c = t.coords[i]
    return t.data[t.offsets[i]:t.offsets[i] + t.block_size(c)].view(t.block_shape(c))


def from_pieces(axes, pieces: Dict[Tuple[int, ...], torch.Tensor], dtype, device) -> BlockTensor:
    keys = sorted(pieces)
    if keys:
        flat = torch.cat([pieces[k].reshape(-1) for k in keys]).to(dtype)
    else:
        flat = torch.zeros(1, dtype=dtype, device=device)
    return BlockTensor(axes, keys, flat)


# ---------------------------------------------------------------------------
# block SVD with truncation
# ---------------------------------------------------------------------------
@dataclass
class SVDResult:
    u: BlockTensor
    s: BlockTensor
    vt: BlockTensor
    trunc_err: float
    D: int
    entropy: float


def _sector_svds(mats) -> dict:
    """Thin SVDs of the per-sector matrices through the eigen-decomposition of
    the Gram matrix on each matrix's smaller side, batched on the GPU by size
    bucket (rocSOLVER's batched symmetric eigensolver; a per-matrix loop is
    latency bound for the ~100 small sectors of a D=2000 state).  Padded rows
    and columns carry eigenvalue -1 so they separate from the non-negative
    spectrum.  Returns q -> (U, S, Vh, rks, cks, ro, co, g) with S descending;
    singular vectors of (near-)zero singular values are completed to an
    orthonormal set."""
    buckets: Dict[int, list] = {}
    for q, (M, *_rest) in mats.items():
        n = min(M.shape)
        b = 128
        while b < n:
            b *= 2
        buckets.setdefault(b, []).append(q)
    out = {}
    for b, qs in buckets.items():
        dev, dt = mats[qs[0]][0].device, mats[qs[0]][0].dtype
        G = torch.zeros((len(qs), b, b), dtype=dt, device=dev)
        for j, q in enumerate(qs):
            M = mats[q][0]
            n = min(M.shape)
            G[j, :n, :n] = M @ M.mT.conj() if M.shape[0] <= M.shape[1] else M.mT.conj() @ M
            if n < b:
                G[j, n:, n:] = -torch.eye(b - n, dtype=dt, device=dev)
        w, V = torch.linalg.eigh(G)
        for j, q in enumerate(qs):
            M, rks, cks, ro, co, g = mats[q]
            n = min(M.shape)
            lam = w[j, b - n:].flip(0).clamp_min(0.0)           # descending
            vec = V[j, :n, b - n:].flip(1)
            S = lam.sqrt().to(torch.float64)
            tiny = S <= 1e-14 * max(float(S[0]), 1e-300)
            Sinv = torch.where(tiny, torch.zeros_like(S), 1.0 / torch.where(tiny, torch.ones_like(S), S))
            if M.shape[0] <= M.shape[1]:
                U = vec
                Vh = (U.mT.conj() @ M) * Sinv[:, None].to(dt)
                if bool(tiny.any()):
                    Vh = _complete_rows(Vh, tiny)
            else:
                Vm = vec
                U = (M @ Vm) * Sinv[None, :].to(dt)
                if bool(tiny.any()):
                    U = _complete_rows(U.mT, tiny).mT
                Vh = Vm.mT.conj()
            out[q] = (U, S, Vh, rks, cks, ro, co, g)
    return out


def _complete_rows(A: torch.Tensor, bad: torch.Tensor) -> torch.Tensor:
    """Replace the rows of A flagged in ``bad`` by unit vectors orthogonal to the others."""
    k, n = A.shape
    good = A[~bad]
    gen = torch.Generator(device="cpu").manual_seed(1234)
    R = torch.randn(int(bad.sum()), n, generator=gen, dtype=torch.float64).to(A.device, A.dtype)
    Q, _ = torch.linalg.qr(torch.cat([good, R]).mT)
    out = A.clone()
    out[bad] = Q.mT[good.shape[0]:]
    return out


def block_svd(t: BlockTensor, ldims: int, ldiv, trunc_err: float, Dmin: int, Dmax: int) -> SVDResult:
    left, right = t.axes[:ldims], t.axes[ldims:]
    ldiv = tuple(ldiv)
    dev, dt = t.data.device, t.data.dtype
    groups: Dict[tuple, dict] = {}
    for i, c in enumerate(t.coords):
        q = ldiv
        for ax, s in zip(left, c[:ldims]):
            sgn = -1 if ax.dir == OUT else 1
            q = tuple(a + sgn * b for a, b in zip(q, ax.qns[s]))
        g = groups.setdefault(q, {"rows": {}, "cols": {}, "blocks": []})
        lk, rk = c[:ldims], c[ldims:]
        g["rows"].setdefault(lk, int(np.prod([ax.degs[s] for ax, s in zip(left, lk)])))
        g["cols"].setdefault(rk, int(np.prod([ax.degs[s] for ax, s in zip(right, rk)])))
        g["blocks"].append((i, lk, rk))
    mats = {}
    for q in sorted(groups):
        g = groups[q]
        rks, cks = sorted(g["rows"]), sorted(g["cols"])
        ro, co = {}, {}
        acc = 0
        for k in rks:
            ro[k] = acc
            acc += g["rows"][k]
        R = acc
        acc = 0
        for k in cks:
            co[k] = acc
            acc += g["cols"][k]
        Cn = acc
        M = torch.zeros((R, Cn), dtype=dt, device=dev)
        for i, lk, rk in g["blocks"]:
            M[ro[lk]:ro[lk] + g["rows"][lk], co[rk]:co[rk] + g["cols"][rk]] = \
                block_view(t, i).reshape(g["rows"][lk], g["cols"][rk])
        mats[q] = (M, rks, cks, ro, co, g)
    dec = _sector_svds(mats)
    svals = [dec[q][1] for q in sorted(groups)]
    sv = torch.cat(svals).double().cpu().numpy()
    owner = np.concatenate([[qi] * len(dec[q][1]) for qi, q in enumerate(sorted(groups))])
    order = np.argsort(-sv, kind="stable")
    sq = sv[order] ** 2
    total = float(sq.sum())
    tail = np.concatenate([np.cumsum(sq[::-1])[::-1], [0.0]])
    D = len(sv)
    for k in range(len(sv) + 1):
        if tail[k] / total <= trunc_err:
            D = k
            break
    D = min(max(D, min(Dmin, len(sv))), Dmax, len(sv))
    D = max(D, 1)
    err = float(tail[D] / total)
    qs_sorted = sorted(groups)
    keep = {}
    for idx in order[:D]:
        q = qs_sorted[owner[idx]]
        keep[q] = keep.get(q, 0) + 1
    kept_qs = [q for q in qs_sorted if q in keep]
    mid_out = Axis.make([(q, keep[q]) for q in kept_qs], OUT)
    mid_in = mid_out.inverse()
    up, sp, vp = {}, {}, {}
    for mi, q in enumerate(kept_qs):
        U, S, Vh, rks, cks, ro, co, g = dec[q]
        k = keep[q]
        sp[(mi, mi)] = torch.diag(S[:k]).to(dt)
        for lk in rks:
            shp = tuple(ax.degs[s] for ax, s in zip(left, lk)) + (k,)
            up[lk + (mi,)] = U[ro[lk]:ro[lk] + g["rows"][lk], :k].reshape(shp)
        for rk in cks:
            shp = (k,) + tuple(ax.degs[s] for ax, s in zip(right, rk))
            vp[(mi,) + rk] = Vh[:k, co[rk]:co[rk] + g["cols"][rk]].reshape(shp)
    u = from_pieces(list(left) + [mid_out], up, dt, dev)
    s = from_pieces([mid_in, mid_out], sp, dt, dev)
    vt = from_pieces([mid_in] + list(right), vp, dt, dev)
    p = np.sort(sv)[::-1][:D] ** 2
    p = p / p.sum()
    p = p[p > 0]
    ee = float(-(p * np.log(p)).sum())
    return SVDResult(u, s, vt, err, D, ee)


# ---------------------------------------------------------------------------
# MPS construction
# ---------------------------------------------------------------------------
def direct_state_mps(m: Model, stat_labs: Sequence[int], device="cuda", dtype=torch.float64) -> List[BlockTensor]:
    """``finite_mps_init.h:213-262``: product state, site 0 carries the total charge."""
    qns, n = m.qns, m.n
    pout = physical_axis(qns, OUT)
    zero = tuple(0 for _ in qns[0])
    total = zero
    for i in range(n):
        total = tuple(a + b for a, b in zip(total, qns[stat_labs[i]]))
    mps = []
    lvb = Axis.make([(zero, 1)], IN)
    lq = zero
    for i in range(n):
        if i == 0:
            rq = tuple(a - b for a, b in zip(total, qns[stat_labs[0]]))
        elif i < n - 1:
            rq = tuple(a - b for a, b in zip(lq, qns[stat_labs[i]]))
        else:
            rq = zero
        rvb = Axis.make([(rq, 1)], OUT)
        blk = np.ones((1, 1, 1))
        mps.append(BlockTensor.from_blocks([lvb, pout, rvb], {(0, stat_labs[i], 0): blk}, device=device,
                                           dtype=dtype))
        lvb = rvb.inverse()
        lq = rq
    return mps


def _charge_counts(qns, nsites):
    cnt = {tuple(0 for _ in qns[0]): 1}
    for _ in range(nsites):
        new = {}
        for q, c in cnt.items():
            for s in qns:
                k = tuple(a + b for a, b in zip(q, s))
                new[k] = new.get(k, 0) + c
        cnt = new
    return cnt


def random_mps(m: Model, total, D: int, seed: int, device="cuda", dtype=torch.float64) -> List[BlockTensor]:
    """Seeded random MPS in sector ``total`` (bond dimension <= D), right-canonical
    from site 1, centre at site 0 (the same construction as the oracle's)."""
    qns, n = m.qns, m.n
    total = tuple(total)
    rng = np.random.default_rng(seed)
    pout = physical_axis(qns, OUT)
    zero = tuple(0 for _ in qns[0])
    bonds = []
    for b in range(n - 1):
        cl, cr = _charge_counts(qns, b + 1), _charge_counts(qns, n - b - 1)
        degs = {}
        for x, c in cl.items():
            q = tuple(a - y for a, y in zip(total, x))
            if q in cr:
                degs[q] = min(c, cr[q])
        tot = sum(degs.values())
        if tot > D:
            degs = {q: max(1, int(round(v * D / tot))) for q, v in degs.items()}
            while sum(degs.values()) > D:
                qm = max(degs, key=lambda q: degs[q])
                degs[qm] -= 1
        bonds.append(Axis.make(sorted(degs.items()), OUT))
    mps = []
    lvb = Axis.make([(zero, 1)], IN)
    for i in range(n):
        rvb = bonds[i] if i < n - 1 else Axis.make([(zero, 1)], OUT)
        div = total if i == 0 else zero
        axes = [lvb, pout, rvb]
        blocks = {}
        for a in range(lvb.nsect):
            for s in range(pout.nsect):
                for c in range(rvb.nsect):
                    if _flow(axes, (a, s, c)) == div:
                        shp = (lvb.degs[a], pout.degs[s], rvb.degs[c])
                        blocks[(a, s, c)] = rng.uniform(-1, 1, size=shp)
        mps.append(BlockTensor.from_blocks(axes, blocks, device=device, dtype=dtype))
        lvb = rvb.inverse()
    for i in range(n - 1, 0, -1):
        r = block_svd(mps[i], 1, divergence(mps[i]), 0.0, 1, 10 ** 9)
        mps[i] = r.vt
        us = contract(r.u, r.s, 1)
        mps[i - 1] = contract(mps[i - 1], us, 1)
    nrm = torch.linalg.vector_norm(mps[0].data).item()
    mps[0].data /= nrm
    return mps


def trivial_env(mps_bond: Axis, mpo_bond: Axis, device, dtype) -> BlockTensor:
    """``vmps_init.h:160-164``."""
    return BlockTensor.from_blocks([mps_bond, mpo_bond.inverse(), mps_bond.inverse()],
                                   {(0, 0, 0): np.ones((1, 1, 1))}, device=device, dtype=dtype)


# ---------------------------------------------------------------------------
# two-site VMPS
# ---------------------------------------------------------------------------
@dataclass
class SweepParams:
    """``algorithm/finite_vmps_sweep_params.h:31-49``."""
    sweeps: int
    Dmin: int
    Dmax: int
    trunc_err: float
    lancz: LanczosParams
    precision: str = "fp64"
    verbose: bool = True


@dataclass
class SweepStats:
    log: list = field(default_factory=list)
    sweep_seconds: list = field(default_factory=list)
    lanczos_s: float = 0.0
    svd_s: float = 0.0
    env_s: float = 0.0
    contract_s: float = 0.0
    plan_s: float = 0.0
    matvecs: int = 0
    algo_flops: float = 0.0


class TwoSiteVMPS:
    """Two-site finite DMRG with GPU-resident MPS and environments."""

    def __init__(self, mps: List[BlockTensor], mpo: List[BlockTensor], params: SweepParams):
        self.mps, self.mpo, self.p = mps, mpo, params
        n = len(mps)
        dev, dt = mps[0].data.device, mps[0].data.dtype
        self.renvs: List[Optional[BlockTensor]] = [None] * n
        self.lenvs: List[Optional[BlockTensor]] = [None] * n
        self.renvs[0] = trivial_env(mps[-1].axes[2], mpo[-1].axes[3], dev, dt)
        for i in range(1, n - 1):
            self.renvs[i] = env_update(self.renvs[i - 1], mps[n - i], mpo[n - i], "r")
        self.lenvs[0] = trivial_env(mps[0].axes[0], mpo[0].axes[0], dev, dt)
        self.stats = SweepStats()

    def update(self, direction: str, target: int) -> float:
        n = len(self.mps)
        p, st = self.p, self.stats
        t0 = time.perf_counter()
        if direction == "r":
            l, r, llen, rlen = target, target + 1, target, n - (target + 2)
        else:
            l, r, llen, rlen = target - 1, target, target - 1, n - target - 1
        t = time.perf_counter()
        psi = contract(self.mps[l], self.mps[r], 1)
        st.contract_s += time.perf_counter() - t
        t = time.perf_counter()
        H = EffectiveHamiltonian([self.lenvs[llen], self.mpo[l], self.mpo[r], self.renvs[rlen]], psi,
                                 precision=p.precision)
        tp = time.perf_counter()
        st.plan_s += tp - t
        res = H.lanczos(psi, p.lancz)
        t_l = time.perf_counter() - tp
        st.lanczos_s += t_l
        st.matvecs += res.iters
        st.algo_flops += res.iters * H.flops()[0]
        H.close()
        t = time.perf_counter()
        sv = block_svd(res.gs_vec, 2, divergence(self.mps[l]), p.trunc_err, p.Dmin, p.Dmax)
        st.svd_s += time.perf_counter() - t
        t = time.perf_counter()
        if direction == "r":
            self.mps[l] = sv.u
            self.mps[r] = contract(sv.s, sv.vt, 1)
        else:
            self.mps[l] = contract(sv.u, sv.s, 1)
            self.mps[r] = sv.vt
        st.contract_s += time.perf_counter() - t
        t = time.perf_counter()
        if direction == "r":
            self.lenvs[llen + 1] = env_update(self.lenvs[llen], self.mps[target], self.mpo[target], "l")
        else:
            self.renvs[rlen + 1] = env_update(self.renvs[rlen], self.mps[target], self.mpo[target], "r")
        st.env_s += time.perf_counter() - t
        tot = time.perf_counter() - t0
        st.log.append(dict(site=target, dir=direction, e0=res.gs_eng, trunc_err=sv.trunc_err, D=sv.D,
                           iters=res.iters, lanczos_s=t_l, total_s=tot, entropy=sv.entropy))
        if p.verbose:
            print(f"Site {target:4d} E0 = {res.gs_eng:20.16f} TruncErr = {sv.trunc_err:.2e} D = {sv.D:5d} "
                  f"Iter = {res.iters:3d} LanczT = {t_l:8.4f} TotT = {tot:8.4f} S = {sv.entropy:10.7f}",
                  file=sys.stderr, flush=True)
        return res.gs_eng

    def sweep(self) -> float:
        n = len(self.mps)
        e0 = None
        t0 = time.perf_counter()
        for i in range(0, n - 2):
            e0 = self.update("r", i)
        for i in range(n - 1, 1, -1):
            e0 = self.update("l", i)
        torch.cuda.synchronize()
        self.stats.sweep_seconds.append(time.perf_counter() - t0)
        return e0

    def run(self) -> float:
        e0 = None
        for _ in range(self.p.sweeps):
            e0 = self.sweep()
        return e0


def TwoSiteFiniteVMPS(mps, mpo, params: SweepParams):
    """Run ``params.sweeps`` sweeps; returns (E0, driver)."""
    drv = TwoSiteVMPS(mps, mpo, params)
    return drv.run(), drv
