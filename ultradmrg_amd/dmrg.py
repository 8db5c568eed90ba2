"""GPU-resident two-site finite DMRG in the reference's DMRG form: the
``FiniteDMRG`` / ``DMRGExecutor`` path that the published Hubbard D=2000
benchmark runs (``include/qlmps/algorithm/dmrg/dmrg_impl.h:140-341``,
``dmrg_init.h:117-309``).

Differences from the VMPS form (``sweep.py``) are in the interface, as in the
reference:

* the Hamiltonian is a ``MatReprMPO``: per site a ``SparMat`` of rank-2 site
  operators W(i, j)[s_in (IN), s_out (OUT)] (``one_dim_tn/mpo/mat_repr_mpo.h:21``,
  ``mpogen/symb_alg/sparse_mat.h:20-80``);
* the environments are block-operator groups, one rank-2 operator per MPO
  bond element, L_i[a (OUT), a' (IN)] and R_k[b (IN), b' (OUT)], updated by
  ``UpdateLeftBlockOps`` / ``UpdateRightBlockOps`` (``dmrg_init.h:131-205``);
* every solve runs on the super-block terms of ``SetEffectiveHamiltonianTerms_``
  (``dmrg_impl.h:319-341``) through ``LanczosSolver(SuperBlockHamiltonianTerms&, ...)``
  (``lanczos_dmrg_solver_impl.h:61-283``).

Device layout.  A block-operator group is held as ONE stacked tensor
env[x_bra, w, x_ket] (the MPO-form environment of the same bond, w running
over the bond elements grouped by charge), so its update is a single
environment contraction on the grouped-GEMM planner instead of a loop of
per-element products and sums; the per-element operators the super-block
terms need are zero-copy views into a transposed copy [w, x_ket, x_bra], in
which each operator's blocks are contiguous.  The site operators of a
MatReprMPO are stacked once into MPO-form tensors for those updates (bond
charges follow from the operators' divergences, left to right).  Nothing is
written to disk: the reference streams block operators through files to
bound host memory (``operator_io.h``); here they stay in HBM.

Boundaries: sweeps run over every bond (left_boundary 0, right_boundary N-1);
the reference's CheckAndUpdateBoundaryMPSTensors skips end sites whose bond
already spans the whole Hilbert space, where an update is an exact SVD and
leaves the state unchanged.
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .solver import EffectiveHamiltonian
from .sweep import SweepParams, SweepStats, block_svd, divergence
from .tensor import IN, OUT, Axis, BlockTensor, contract, env_update, transpose


# ---------------------------------------------------------------------------
# MatReprMPO
# ---------------------------------------------------------------------------
@dataclass
class SparMat:
    """``sparse_mat.h:20-80``: rows x cols site operators, null where absent."""
    rows: int
    cols: int
    data: Dict[Tuple[int, int], BlockTensor] = field(default_factory=dict)

    def IsNull(self, i: int, j: int) -> bool:
        return (i, j) not in self.data

    def __call__(self, i: int, j: int) -> BlockTensor:
        return self.data[(i, j)]

    def SetElem(self, i: int, j: int, op: BlockTensor):
        self.data[(i, j)] = op


def mat_repr_mpo(mpo: Sequence[BlockTensor]) -> List[SparMat]:
    """The MatReprMPO of MPO-form tensors W[w_l, s_in, s_out, w_r]: entry
    (i, j) is the rank-2 operator W[i, :, :, j] over the bond elements i, j
    (sector-major order), present where it has a nonzero value."""
    out = []
    for W in mpo:
        lax, rax = W.axes[0], W.axes[3]
        loff = np.concatenate([[0], np.cumsum(lax.degs)])
        roff = np.concatenate([[0], np.cumsum(rax.degs)])
        per: Dict[Tuple[int, int], Dict[Tuple[int, int], np.ndarray]] = {}
        for (ws, si, so, wr), blk in W.blocks().items():
            for a in range(blk.shape[0]):
                for z in range(blk.shape[3]):
                    v = blk[a, :, :, z]
                    if np.any(v != 0):
                        per.setdefault((int(loff[ws] + a), int(roff[wr] + z)), {})[(si, so)] = v
        sm = SparMat(lax.dim, rax.dim)
        for (i, j), blocks in per.items():
            sm.data[(i, j)] = BlockTensor.from_blocks([W.axes[1], W.axes[2]], blocks, device=W.data.device,
                                                      dtype=W.data.dtype)
        out.append(sm)
    return out


class _Bond:
    """An MPO bond's elements grouped by charge: ``axis`` has one sector per
    distinct charge (ascending); element e sits at ``where[e]`` = (sector,
    position)."""

    def __init__(self, qns: Sequence[Tuple[int, ...]], dir: int):
        groups: Dict[Tuple[int, ...], List[int]] = {}
        for e, q in enumerate(qns):
            groups.setdefault(tuple(q), []).append(e)
        self.where: List[Tuple[int, int]] = [(0, 0)] * len(qns)
        secs = []
        for s, q in enumerate(sorted(groups)):
            for p, e in enumerate(groups[q]):
                self.where[e] = (s, p)
            secs.append((q, len(groups[q])))
        self.axis = Axis.make(secs, dir)


def _div(op: BlockTensor) -> Tuple[int, ...]:
    c = op.coords[0]
    q = None
    for ax, s in zip(op.axes, c):
        t = ax.qns[s] if ax.dir == OUT else tuple(-v for v in ax.qns[s])
        q = t if q is None else tuple(a + b for a, b in zip(q, t))
    return q


def stack_mat_repr(mro: Sequence[SparMat]) -> Tuple[List[BlockTensor], List[_Bond]]:
    """MPO-form tensors W[w_l (IN), s_in, s_out, w_r (OUT)] of a MatReprMPO and
    its bonds (bonds[k] lies left of site k; bonds[N] right of the last).
    Element charges: the left end's single element has charge 0 and
    q(w_r) = q(w_l) - div(W(i, j)), so every stacked tensor has divergence 0."""
    n = len(mro)
    qn_len = None
    for sm in mro:
        for op in sm.data.values():
            if op.coords:
                qn_len = len(op.axes[0].qns[0])
                break
        if qn_len:
            break
    if qn_len is None:
        raise ValueError("MatReprMPO without any operator")
    zero = tuple(0 for _ in range(qn_len))
    lq: List[Optional[Tuple[int, ...]]] = [zero] * mro[0].rows
    Ws, bonds = [], [_Bond(lq, IN)]
    for k, sm in enumerate(mro):
        if len(lq) != sm.rows:
            raise ValueError(f"site {k}: {sm.rows} rows, previous site has {len(lq)} columns")
        rq: List[Optional[Tuple[int, ...]]] = [None] * sm.cols
        phys = None
        for (i, j), op in sorted(sm.data.items()):
            if not op.coords:
                continue
            phys = phys or (op.axes[0], op.axes[1])
            q = tuple(a - b for a, b in zip(lq[i], _div(op)))
            if rq[j] is None:
                rq[j] = q
            elif rq[j] != q:
                raise ValueError(f"site {k}: column {j} has operators of different charge")
        if phys is None:
            raise ValueError(f"site {k}: no operator")
        rq = [q if q is not None else zero for q in rq]   # columns without operators: any charge
        lb, rb = bonds[-1], _Bond(rq, OUT)
        lax, rax = Axis(lb.axis.qns, lb.axis.degs, IN), rb.axis
        blocks: Dict[Tuple[int, int, int, int], np.ndarray] = {}
        for (i, j), op in sm.data.items():
            ls, lp = lb.where[i]
            rs, rp = rb.where[j]
            for (si, so), v in op.blocks().items():
                key = (ls, si, so, rs)
                if key not in blocks:
                    blocks[key] = np.zeros((lax.degs[ls], phys[0].degs[si], phys[1].degs[so], rax.degs[rs]),
                                           dtype=v.dtype)
                blocks[key][lp, :, :, rp] = v
        dev = next(iter(sm.data.values())).data.device
        dt = next(iter(sm.data.values())).data.dtype
        Ws.append(BlockTensor.from_blocks([lax, phys[0], phys[1], rax], blocks, device=dev, dtype=dt))
        bonds.append(rb)
        lq = rq
    return Ws, bonds


# ---------------------------------------------------------------------------
# block-operator groups
# ---------------------------------------------------------------------------
class BlockOperatorGroup:
    """``LeftBlockOperatorGroup`` / ``RightBlockOperatorGroup``
    (``algorithm/dmrg/dmrg.h:17-20``): element e of the MPO bond -> a rank-2
    block operator [x_ket, x_bra], stored stacked as env[x_bra, w, x_ket]."""

    def __init__(self, env: BlockTensor, bond: _Bond):
        self.env, self.bond = env, bond
        self._ops: Optional[List[Optional[BlockTensor]]] = None
        self._t = None

    def __len__(self):
        return len(self.bond.where)

    def release_views(self):
        """Drop the transposed copy and the views (the stack itself stays)."""
        self._ops = None
        self._t = None

    def ops(self) -> List[Optional[BlockTensor]]:
        """Per element, the operator as a zero-copy view into one transposed
        copy of the stack (None for an element without blocks)."""
        if self._ops is None:
            t = transpose(self.env, (1, 2, 0))     # [w, x_ket, x_bra]
            self._t = t
            xk, xb = t.axes[1], t.axes[2]
            by_sect: Dict[int, List[Tuple[int, int, int]]] = {}
            for c, o in zip(t.coords, t.offsets):
                by_sect.setdefault(c[0], []).append((c[1], c[2], o))
            ops: List[Optional[BlockTensor]] = []
            for s, p in self.bond.where:
                blks = by_sect.get(s, [])
                if not blks:
                    ops.append(None)
                    continue
                coords = [(a, b) for a, b, _ in blks]
                offs = [o + p * xk.degs[a] * xb.degs[b] for a, b, o in blks]
                ops.append(BlockTensor([xk, xb], coords, t.data, offs))
            self._ops = ops
        return self._ops


def _trivial_group(mps_bond: Axis, mpo_bond: Axis, device, dtype) -> BlockTensor:
    """The 1 x 1 identity of ``dmrg_init.h:220-225, 247-252`` in stacked form
    (mpo_bond: the end axis of the end site's stacked operator tensor)."""
    return BlockTensor.from_blocks([mps_bond, mpo_bond.inverse(), mps_bond.inverse()],
                                   {(0, 0, 0): np.ones((1, 1, 1))}, device=device, dtype=dtype)


def UpdateLeftBlockOps(log: BlockOperatorGroup, mps: BlockTensor, W: BlockTensor,
                       right_bond: _Bond) -> BlockOperatorGroup:
    """``dmrg_init.h:179-205``: log_next[col] = A^dag (sum_row L_row (x) W(row, col)) A,
    as one environment contraction of the stack with the stacked site operators."""
    return BlockOperatorGroup(env_update(log.env, mps, W, "l"), right_bond)


def UpdateRightBlockOps(rog: BlockOperatorGroup, mps: BlockTensor, W: BlockTensor,
                        left_bond: _Bond) -> BlockOperatorGroup:
    """``dmrg_init.h:131-161``: rog_next[row] = B (sum_col W(row, col) (x) R_col) B^dag."""
    return BlockOperatorGroup(env_update(rog.env, mps, W, "r"), left_bond)


# ---------------------------------------------------------------------------
# the executor
# ---------------------------------------------------------------------------
class DMRGExecutor:
    """``DMRGExecutor`` (``dmrg_impl.h:63-341``) on a right-canonical MPS."""

    def __init__(self, mps: List[BlockTensor], mro: Sequence[SparMat], params: SweepParams):
        self.mps, self.mro, self.p = mps, list(mro), params
        self.N = len(mps)
        if len(self.mro) != self.N:
            raise ValueError("MatReprMPO and MPS differ in length")
        self.W, self.bonds = stack_mat_repr(self.mro)
        self.stats = SweepStats()
        self.sweeps_done = 0
        self._flags = None
        self.e0 = None
        self.DMRGInit_()

    def DMRGInit_(self):
        """Right block-operator groups of every length and the left boundary
        group (``InitBlockOps`` / ``UpdateBoundaryBlockOps``, ``dmrg_init.h:207-309``)."""
        n = self.N
        dev, dt = self.mps[0].data.device, self.mps[0].data.dtype
        self.ropg: List[Optional[BlockOperatorGroup]] = [None] * n
        self.lopg: List[Optional[BlockOperatorGroup]] = [None] * n
        self.ropg[0] = BlockOperatorGroup(_trivial_group(self.mps[-1].axes[2], self.W[-1].axes[3], dev, dt),
                                          self.bonds[n])
        for i in range(1, n - 1):
            self.ropg[i] = UpdateRightBlockOps(self.ropg[i - 1], self.mps[n - i], self.W[n - i], self.bonds[n - i])
        self.lopg[0] = BlockOperatorGroup(_trivial_group(self.mps[0].axes[0], self.W[0].axes[0], dev, dt),
                                          self.bonds[0])

    def SetEffectiveHamiltonianTerms_(self, l: int, r: int):
        """``dmrg_impl.h:319-341``: per middle bond element j, the block-site
        terms (L_i, W_l(i, j)) and site-block terms (W_r(j, k), R_k).  Terms
        whose block operator has no blocks (a zero operator) and groups with
        no term on one side contribute nothing and are left out."""
        lops = self.lopg[l].ops()
        rops = self.ropg[self.N - 1 - r].ops()
        sl, sr = self.mro[l], self.mro[r]
        terms = []
        for j in range(sl.cols):
            bs = [(lops[i], sl(i, j)) for i in range(len(lops)) if not sl.IsNull(i, j) and lops[i] is not None]
            sb = [(sr(j, k), rops[k]) for k in range(sr.cols) if not sr.IsNull(j, k) and rops[k] is not None]
            if bs and sb:
                terms.append((bs, sb))
        return terms

    def TwoSiteUpdate_(self, direction: str, l: int, r: int) -> float:
        """``dmrg_impl.h:213-317``."""
        p, st, n = self.p, self.stats, self.N
        t0 = time.perf_counter()
        llen, rlen = l, n - 1 - r
        t = time.perf_counter()
        psi = contract(self.mps[l], self.mps[r], 1)
        st.contract_s += time.perf_counter() - t
        t = time.perf_counter()
        terms = self.SetEffectiveHamiltonianTerms_(l, r)
        H = EffectiveHamiltonian(terms, psi, precision=p.precision)
        # the plan holds its own stacked operators: the views can go
        self.lopg[llen].release_views()
        self.ropg[rlen].release_views()
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
            self.ropg[rlen] = None
            self.lopg[llen + 1] = UpdateLeftBlockOps(self.lopg[llen], self.mps[l], self.W[l], self.bonds[l + 1])
        else:
            self.lopg[llen] = None
            self.ropg[rlen + 1] = UpdateRightBlockOps(self.ropg[rlen], self.mps[r], self.W[r], self.bonds[r])
        st.env_s += time.perf_counter() - t
        tot = time.perf_counter() - t0
        st.log.append(dict(l=l, r=r, dir=direction, e0=res.gs_eng, trunc_err=sv.trunc_err, D=sv.D,
                           iters=res.iters, lanczos_s=t_l, total_s=tot, entropy=sv.entropy))
        if p.verbose:
            print(f"Site ({l:4d},{r:4d}) E0 = {res.gs_eng:20.16f} TruncErr = {sv.trunc_err:.2e} D = {sv.D:5d} "
                  f"Iter = {res.iters:3d} LanczT = {t_l:8.4f} TotT = {tot:8.4f} S = {sv.entropy:10.7f}",
                  file=sys.stderr, flush=True)
        return res.gs_eng

    def DMRGSweep_(self) -> Optional[float]:
        """``dmrg_impl.h:172-202``; None after an emergency stop (state dumped)."""
        n = self.N
        t0 = time.perf_counter()
        for direction, pairs in (("r", [(i, i + 1) for i in range(0, n - 2)]),
                                 ("l", [(i - 1, i) for i in range(n - 1, 1, -1)])):
            for l, r in pairs:
                self.e0 = self.TwoSiteUpdate_(direction, l, r)
                if self._flags is not None and self._flags.emergency:
                    from .checkpoint import save_mps
                    save_mps(self.p.checkpoint_dir, self.mps, {"complete_sweep": False,
                                                               "sweeps_done": self.sweeps_done,
                                                               "stopped_at": [direction, l], "energy": self.e0})
                    return None
        torch.cuda.synchronize()
        self.stats.sweep_seconds.append(time.perf_counter() - t0)
        return self.e0

    def sweep(self) -> Optional[float]:
        return self.DMRGSweep_()

    def Execute(self) -> float:
        """``dmrg_impl.h:140-165``: sweeps until ``params.sweeps`` are done; with
        ``checkpoint_dir`` set, the MPS is dumped after every sweep and the
        stop signals act as in the VMPS driver (``dmrg_impl.h:478-502``)."""
        from .checkpoint import StopFlags, save_mps
        flags = StopFlags().install() if self.p.checkpoint_dir else None
        self._flags = flags
        e0 = None
        try:
            for _ in range(self.sweeps_done, self.p.sweeps):
                e0 = self.DMRGSweep_()
                if e0 is None:
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


def FiniteDMRG(mps: List[BlockTensor], mro: Sequence[SparMat], params: SweepParams):
    """``FiniteDMRG`` (``dmrg_impl.h:504-511``): returns (E0, executor)."""
    ex = DMRGExecutor(mps, mro, params)
    return ex.Execute(), ex
