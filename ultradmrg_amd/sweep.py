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
block SVD through ``qlg_block_svd``.  The truncation rule is the library's:
pooled singular values, smallest D whose discarded weight
sum(s^2)/total <= trunc_err, clamped to [Dmin, Dmax].
"""
from __future__ import annotations

import ctypes as C
import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .mpo import Model, physical_axis
from .solver import EffectiveHamiltonian, LanczosParams
from . import _lib as L
from .tensor import IN, OUT, Axis, BlockTensor, _flow, contract, default_context, env_update


def divergence(t: BlockTensor) -> Tuple[int, ...]:
    return _flow(t.axes, t.coords[0])


def block_view(t: BlockTensor, i: int) -> torch.Tensor:
    c = t.coords[i]
    return t.data[t.offsets[i]:t.offsets[i] + t.block_size(c)].view(t.block_shape(c))


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


def eigh_batched(mats: List[torch.Tensor]) -> List[Tuple[torch.Tensor, torch.Tensor]]:
    """Eigen-decompositions of batches of Hermitian matrices ([B, n, n] each, all
    of one dtype, n > 1) through ``qlg_eigh_batched``: one library call for all
    batches.  Returns (w [B, n] ascending, V [B, n, n] with eigenvectors as
    columns).  The inputs are consumed (they hold the eigenvectors on return)."""
    if not mats:
        return []
    dev, dt = mats[0].device, mats[0].dtype
    ctx = default_context(dev.index or 0)
    groups = (L.EighGroup * len(mats))()
    ws = []
    for g, A in zip(groups, mats):
        assert A.is_contiguous() and A.dim() == 3 and A.shape[1] == A.shape[2] and A.dtype == dt
        w = torch.empty(A.shape[0], A.shape[1], dtype=torch.float64, device=dev)
        g.n, g.batch, g.a, g.w = A.shape[1], A.shape[0], A.data_ptr(), w.data_ptr()
        ws.append(w)
    torch.cuda.synchronize()     # the library works on its own streams; it returns after the work is done
    ctx.check(ctx.lib.qlg_eigh_batched(ctx.h, L.QLG_C128 if dt.is_complex else L.QLG_F64, len(mats), groups))
    # row j of each buffer is eigenvector j (conjugated for complex data)
    return [(w, A.mT.conj()) for w, A in zip(ws, mats)]


def block_svd(t: BlockTensor, ldims: int, ldiv, trunc_err: float, Dmin: int, Dmax: int) -> SVDResult:
    """The reference's SVD(t, ldims, div, trunc_err, Dmin, Dmax, u, s, vt,
    actual_trunc_err, D) (two_site_update_finite_vmps_impl.h:302-307) through
    ``qlg_block_svd``: sector matrices gathered on the GPU, FP64 Gram
    eigen-decompositions, back-projection, refinement of kept values below
    1e-6 of their sector's largest, Newton-Schulz orthonormalisation and the
    assembly of u, s and vt, all in the library.  Truncation (the library's
    rule): pooled singular values, smallest D whose discarded weight
    sum(s^2)/total <= trunc_err, clamped to [Dmin, Dmax].  s is returned in
    t's dtype (its values are real); ``entropy`` is the von Neumann entropy of
    the kept weights."""
    dev = t.data.device
    ctx = default_context(dev.index or 0)
    lib = ctx.lib
    qlen = t.qn_len
    desc = t.descriptor()
    ld = (C.c_int32 * qlen)(*[int(v) for v in ldiv])
    h = C.c_void_p()
    torch.cuda.synchronize()
    ctx.check(lib.qlg_block_svd(ctx.h, C.byref(desc), int(ldims), ld, float(trunc_err), int(Dmin), int(Dmax),
                                C.byref(h)))
    try:
        err, D, ns = C.c_double(), C.c_int64(), C.c_int64()
        ctx.check(lib.qlg_svd_result(h, C.byref(err), C.byref(D), C.byref(ns)))
        qn = (C.c_int32 * max(ns.value * qlen, 1))()
        deg = (C.c_int64 * max(ns.value, 1))()
        ctx.check(lib.qlg_svd_bond(h, qn, deg))
        mid_out = Axis.make([(tuple(qn[i * qlen:(i + 1) * qlen]), deg[i]) for i in range(ns.value)], OUT)
        mid_in = mid_out.inverse()

        def fetch(which, axes, dtype):
            nb, size = C.c_int64(), C.c_int64()
            ctx.check(lib.qlg_svd_tensor(h, which, C.byref(nb), C.byref(size), None, None, None))
            r = len(axes)
            coords = (C.c_int32 * max(nb.value * r, 1))()
            offs = (C.c_int64 * max(nb.value, 1))()
            data = C.c_void_p()
            ctx.check(lib.qlg_svd_tensor(h, which, None, None, coords, offs, C.byref(data)))
            buf = torch.empty(max(size.value, 1), dtype=dtype, device=dev)
            nbytes = size.value * buf.element_size()
            if nbytes:
                ctx.check(lib.qlg_memcpy_d2d(ctx.h, C.c_void_p(buf.data_ptr()), data, nbytes))
            cl = [tuple(coords[b * r:(b + 1) * r]) for b in range(nb.value)]
            return BlockTensor(axes, cl, buf, list(offs[:nb.value]))

        dt = t.data.dtype
        u = fetch(0, list(t.axes[:ldims]) + [mid_out], dt)
        s = fetch(1, [mid_in, mid_out], torch.float64)
        vt = fetch(2, [mid_in] + list(t.axes[ldims:]), dt)
    finally:
        lib.qlg_svd_destroy(h)
    # kept singular values: the diagonals of s's blocks
    diag = torch.cat([s.data[o:o + k * k].view(k, k).diagonal()
                      for (mi, _), o, k in zip(s.coords, s.offsets, mid_out.degs)]) if s.coords else s.data[:0]
    pr = diag.double() ** 2
    pr = pr / pr.sum() if float(pr.sum()) > 0 else pr
    pr = pr[pr > 0]
    ee = float(-(pr * torch.log(pr)).sum()) if pr.numel() else 0.0
    if dt.is_complex:
        s = BlockTensor(s.axes, s.coords, s.data.to(dt), s.offsets)
    return SVDResult(u, s, vt, float(err.value), int(D.value), ee)


# ---------------------------------------------------------------------------
# MPS construction
# ---------------------------------------------------------------------------
def direct_state_mps(m: Model, stat_labs: Sequence[int], device="cuda", dtype=torch.float64) -> List[BlockTensor]:
    """``finite_mps_init.h:213-262``: product state, site 0 carries the total charge."""
    n = m.n
    qat = m.qns_at
    zero = tuple(0 for _ in qat(0)[0])
    total = zero
    for i in range(n):
        total = tuple(a + b for a, b in zip(total, qat(i)[stat_labs[i]]))
    mps = []
    lvb = Axis.make([(zero, 1)], IN)
    lq = zero
    for i in range(n):
        pout = physical_axis(qat(i), OUT)
        if i == 0:
            rq = tuple(a - b for a, b in zip(total, qat(0)[stat_labs[0]]))
        elif i < n - 1:
            rq = tuple(a - b for a, b in zip(lq, qat(i)[stat_labs[i]]))
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


def random_bonds(m: Model, total, D: int) -> List[Axis]:
    """Virtual bond sectors of a random MPS in sector ``total``: every charge a
    bond can carry, with degeneracy min(count from the left, count from the
    right), scaled down to <= D in total."""
    qns, n = m.qns, m.n
    total = tuple(total)
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
    return bonds


def _site_axes(m: Model, total, bonds: List[Axis]):
    """(axes, divergence) of every site tensor for the given bonds."""
    pout = physical_axis(m.qns, OUT)
    zero = tuple(0 for _ in m.qns[0])
    lvb = Axis.make([(zero, 1)], IN)
    out = []
    for i in range(m.n):
        rvb = bonds[i] if i < m.n - 1 else Axis.make([(zero, 1)], OUT)
        out.append(([lvb, pout, rvb], tuple(total) if i == 0 else zero))
        lvb = rvb.inverse()
    return out


def random_mps(m: Model, total, D: int, seed: int, device="cuda", dtype=torch.float64,
               canonical: bool = True) -> List[BlockTensor]:
    """Seeded random MPS in sector ``total`` (bond dimension <= D), right-canonical
    from site 1, centre at site 0 (the same construction as the oracle's).
    ``canonical=False`` returns the raw site tensors, filled with U[-1, 1) by a
    seeded device generator (fast for long chains; not the oracle's stream)."""
    n = m.n
    rng = np.random.default_rng(seed)
    gen = None if canonical else torch.Generator(device=device).manual_seed(seed)
    mps = []
    for axes, div in _site_axes(m, total, random_bonds(m, total, D)):
        lvb, pout, rvb = axes
        blocks = {}
        for a in range(lvb.nsect):
            for s in range(pout.nsect):
                for c in range(rvb.nsect):
                    if _flow(axes, (a, s, c)) == div:
                        shp = (lvb.degs[a], pout.degs[s], rvb.degs[c])
                        blocks[(a, s, c)] = rng.uniform(-1, 1, size=shp) if canonical else shp
        if canonical:
            mps.append(BlockTensor.from_blocks(axes, blocks, device=device, dtype=dtype))
        else:
            keys = sorted(blocks)
            size = sum(int(np.prod(blocks[k])) for k in keys)
            data = torch.rand(max(size, 1), generator=gen, device=device, dtype=dtype) * 2 - 1
            mps.append(BlockTensor(axes, keys, data))
    if not canonical:
        return mps
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
    checkpoint_dir: Optional[str] = None   # MPS dumped here after every sweep (and on a stop signal)


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
        self.sweeps_done = 0
        self._flags = None

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

    def sweep(self) -> Optional[float]:
        """One right and one left half-sweep; returns the last update's energy, or
        None after an emergency stop (the MPS is dumped mid-sweep)."""
        n = len(self.mps)
        e0 = None
        t0 = time.perf_counter()
        for direction, sites in (("r", range(0, n - 2)), ("l", range(n - 1, 1, -1))):
            for i in sites:
                e0 = self.update(direction, i)
                if self._flags is not None and self._flags.emergency:
                    from .checkpoint import save_mps
                    save_mps(self.p.checkpoint_dir, self.mps, {"complete_sweep": False,
                                                               "sweeps_done": self.sweeps_done,
                                                               "stopped_at": [direction, i], "energy": e0})
                    return None
        torch.cuda.synchronize()
        self.stats.sweep_seconds.append(time.perf_counter() - t0)
        return e0

    def run(self) -> float:
        """Sweeps until ``params.sweeps`` are done in total (a resumed driver
        continues from its ``sweeps_done``).  With ``checkpoint_dir`` set, the MPS is dumped
        after every sweep; SIGUSR1 / SIGUSR2 stop after the current sweep and
        SIGINT / SIGTERM after the current update (``dmrg_impl.h:478-502``), in
        both cases with the state dumped first."""
        from .checkpoint import StopFlags, save_mps
        e0 = None
        flags = StopFlags().install() if self.p.checkpoint_dir else None
        self._flags = flags
        try:
            for k in range(self.sweeps_done, self.p.sweeps):
                e0 = self.sweep()
                if e0 is None:        # emergency stop inside the sweep (already dumped)
                    break
                self.sweeps_done += 1
                if self.p.checkpoint_dir:
                    save_mps(self.p.checkpoint_dir, self.mps,
                             {"complete_sweep": True, "sweeps_done": self.sweeps_done, "energy": e0})
                if flags and flags.stop:
                    break
        finally:
            if flags:
                flags.uninstall()
            self._flags = None
        return e0

    @classmethod
    def resume(cls, checkpoint_dir: str, mpo: List[BlockTensor], params: SweepParams, device="cuda"):
        """Continue from a dump.  A dump taken inside a sweep is right-canonicalised
        (centre moved to site 0) first; the state itself is unchanged."""
        from .checkpoint import load_mps
        mps, man = load_mps(checkpoint_dir, device)
        if not man.get("complete_sweep", False):
            center_to_zero(mps)
        drv = cls(mps, mpo, params)
        drv.sweeps_done = int(man.get("sweeps_done", 0))
        return drv


def center_to_zero(mps: List[BlockTensor], start: Optional[int] = None):
    """Right-canonicalise sites start .. 1 (default N-1 .. 1) by SVD without
    truncation: every singular value is kept (Dmin = the full rank), so bonds
    keep their dimension and the state is unchanged."""
    n = len(mps)
    for i in range(n - 1 if start is None else start, 0, -1):
        r = block_svd(mps[i], 1, divergence(mps[i]), 0.0, 10 ** 9, 10 ** 9)
        mps[i] = r.vt
        mps[i - 1] = contract(mps[i - 1], contract(r.u, r.s, 1), 1)


def TwoSiteFiniteVMPS(mps, mpo, params: SweepParams):
    """Run ``params.sweeps`` sweeps; returns (E0, driver)."""
    drv = TwoSiteVMPS(mps, mpo, params)
    return drv.run(), drv
