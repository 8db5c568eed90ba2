"""Python mirror of the reference's effective-Hamiltonian / Krylov call surface,
running on the HIP library.

Reference names kept (``include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h``,
``include/qlmps/algorithm/tdvp/lanczos_expmv_solver_impl.h``,
``include/qlmps/algorithm/lanczos_params.h``):

* ``LanczosParams(error, max_iterations)`` (defaults 1e-7 / 200)
* ``LanczosRes(iters, gs_eng, gs_vec)`` and ``ExpmvRes(iters, expmv)``
* ``LanczosSolver(eff_ham, init_state, matvec, params)`` where ``eff_ham`` is
  ``[L, W_l, W_r, R]`` (or ``[L, W, R]``) and ``matvec`` is
  ``eff_ham_mul_two_site_state`` / ``eff_ham_mul_single_site_state``
* ``LanczosExpmvSolver(eff_ham, init_state, matvec, step_length, params)``
* ``TridiagGsSolver`` / ``TridiagExpme1Solver``

Differences from the reference that a caller sees: the initial state is not
consumed (it is left untouched), and the returned vectors are new device
tensors in the plan's Krylov layout.  ``EffectiveHamiltonian`` holds the plan
(the sector pairing) so that repeated solves with the same operators skip
re-planning.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .tensor import BlockTensor, Context, default_context

PRECISIONS = {"fp64": L.QLG_PREC_F64, "fp32": L.QLG_PREC_F32, "bf16x3": L.QLG_PREC_BF16X3,
              "bf16x6": L.QLG_PREC_BF16X6}


@dataclass
class LanczosParams:
    error: float = 1.0e-7
    max_iterations: int = 200


@dataclass
class LanczosRes:
    iters: int
    gs_eng: float
    gs_vec: BlockTensor


@dataclass
class ExpmvRes:
    iters: int
    expmv: BlockTensor


class EffectiveHamiltonian:
    """A planned H_eff = L . W... . R acting on states with ``state``'s structure."""

    def __init__(self, eff_ham, state: BlockTensor, precision: str = "fp64",
                 ctx: Optional[Context] = None):
        """``eff_ham``: [L, W_l, W_r, R] (two-site), [L, W, R] (one-site), or the
        DMRG form's SuperBlockHamiltonianTerms: a list over the middle MPO bond
        of pairs (block-site terms [(L_i, W_l(i, j)), ...], site-block terms
        [(W_r(j, k), R_k), ...]) (``lanczos_dmrg_solver_impl.h:199-283``)."""
        self.ctx = ctx or default_context(state.data.device.index or 0)
        self.ops = list(eff_ham)          # keep operand buffers alive
        lib = self.ctx.lib
        h = C.c_void_p()
        sd = state.descriptor()
        if self.ops and isinstance(self.ops[0], tuple):
            self._plan_super_block(self.ops, sd, precision, h)
        else:
            self._plan_mpo_form(sd, precision, h)
        self._finish(h, state)

    def _plan_super_block(self, terms, sd, precision, h):
        lib = self.ctx.lib
        descs = {}                        # one descriptor per distinct tensor object

        def dptr(t):
            if id(t) not in descs:
                descs[id(t)] = (t, t.descriptor())
            return C.pointer(descs[id(t)][1])
        nl = (C.c_int64 * len(terms))(*[len(bs) for bs, _ in terms])
        nr = (C.c_int64 * len(terms))(*[len(sb) for _, sb in terms])
        lt = [(L_i, W) for bs, _ in terms for L_i, W in bs]
        rt = [(W, R_k) for _, sb in terms for W, R_k in sb]
        left = (L.Term * max(len(lt), 1))(*[L.Term(dptr(a), dptr(b)) for a, b in lt])
        right = (L.Term * max(len(rt), 1))(*[L.Term(dptr(a), dptr(b)) for a, b in rt])
        torch.cuda.synchronize()
        self.ctx.check(lib.qlg_plan_super_block(self.ctx.h, len(terms), nl, left, nr, right, C.byref(sd),
                                                PRECISIONS[precision], C.byref(h)))

    def _plan_mpo_form(self, sd, precision, h):
        lib = self.ctx.lib
        descs = [t.descriptor() for t in self.ops]
        torch.cuda.synchronize()
        if len(self.ops) == 4:
            rc = lib.qlg_plan_two_site(self.ctx.h, *[C.byref(d) for d in descs], C.byref(sd),
                                       PRECISIONS[precision], C.byref(h))
        elif len(self.ops) == 3:
            rc = lib.qlg_plan_one_site(self.ctx.h, *[C.byref(d) for d in descs], C.byref(sd),
                                       PRECISIONS[precision], C.byref(h))
        else:
            raise ValueError("eff_ham must be [L, W_l, W_r, R] or [L, W, R]")
        self.ctx.check(rc)

    def _finish(self, h, state):
        lib = self.ctx.lib
        self.h = h
        self.lib = lib
        nb, size = C.c_int64(), C.c_int64()
        self.ctx.check(lib.qlg_plan_vector_layout(h, C.byref(nb), C.byref(size), None, None))
        coords = (C.c_int32 * max(nb.value * state.rank, 1))()
        offs = (C.c_int64 * max(nb.value, 1))()
        self.ctx.check(lib.qlg_plan_vector_layout(h, None, None, coords, offs))
        r = state.rank
        self.layout_coords = [tuple(coords[b * r:(b + 1) * r]) for b in range(nb.value)]
        self.layout_offsets = list(offs[:nb.value])
        self.size = size.value
        self.axes = state.axes
        self.dtype = state.data.dtype
        self.device = state.data.device

    # ----- layout helpers ------------------------------------------------------------
    def new_vector(self, zero=True) -> BlockTensor:
        buf = (torch.zeros if zero else torch.empty)(max(self.size, 1), dtype=self.dtype, device=self.device)
        return BlockTensor(self.axes, self.layout_coords, buf, self.layout_offsets)

    def adopt(self, state: BlockTensor) -> BlockTensor:
        """Return ``state`` in the plan's Krylov layout (copying blocks if the layouts differ).
        The copy is one device gather/scatter over all blocks."""
        if state.coords == self.layout_coords and state.offsets == self.layout_offsets:
            return state
        out = self.new_vector()
        pos = {c: o for c, o in zip(self.layout_coords, self.layout_offsets)}
        sizes = [state.block_size(c) for c in state.coords]
        keep = [i for i, n in enumerate(sizes) if n > 0]
        if not keep:
            return out
        dev = out.data.device
        n = torch.tensor([sizes[i] for i in keep], dtype=torch.int64)
        src0 = torch.tensor([state.offsets[i] for i in keep], dtype=torch.int64)
        dst0 = torch.tensor([pos[state.coords[i]] for i in keep], dtype=torch.int64)
        total = int(n.sum())
        prefix = torch.cumsum(n, 0) - n
        table = torch.stack([src0 - prefix, dst0 - prefix]).to(dev)          # per-block shift
        blk = torch.repeat_interleave(torch.arange(len(keep), device=dev), n.to(dev), output_size=total)
        k = torch.arange(total, device=dev)
        out.data[table[1][blk] + k] = state.data[table[0][blk] + k]
        return out

    def flops(self):
        a, e = C.c_double(), C.c_double()
        self.ctx.check(self.lib.qlg_plan_flops(self.h, C.byref(a), C.byref(e)))
        return a.value, e.value

    def workspace_bytes(self) -> int:
        b = C.c_int64()
        self.ctx.check(self.lib.qlg_plan_workspace(self.h, C.byref(b)))
        return b.value

    def set_precision(self, precision: str):
        self.ctx.check(self.lib.qlg_plan_set_precision(self.h, PRECISIONS[precision]))

    def set_reorth(self, passes: int):
        """Full reorthogonalisation of each Lanczos residual against the whole
        basis (classical Gram-Schmidt, ``passes`` = 0, 1 or 2 per iteration)."""
        self.ctx.check(self.lib.qlg_plan_set_reorth(self.h, int(passes)))

    def use_graphs(self, on: bool = True):
        """Capture the Krylov solvers' iteration chunks as hipGraphs on first use
        and replay them in later solves on this plan."""
        self.ctx.check(self.lib.qlg_plan_use_graphs(self.h, int(on)))

    def enable_timing(self, on=True):
        self.ctx.check(self.lib.qlg_plan_enable_timing(self.h, int(on)))

    def stage_flops(self):
        out = (C.c_double * 3)()
        self.ctx.check(self.lib.qlg_plan_stage_flops(self.h, out))
        return list(out)

    def stage_work(self):
        """(algorithmic FLOPs, algorithmic HBM bytes) per stage of one matvec."""
        f, b = (C.c_double * 3)(), (C.c_double * 3)()
        self.ctx.check(self.lib.qlg_plan_stage_work(self.h, f, b))
        return list(f), list(b)

    def stage_ms(self):
        out = (C.c_double * 3)()
        self.ctx.check(self.lib.qlg_plan_stage_ms(self.h, out))
        return list(out)

    # ----- operations -------------------------------------------------------------------
    def matvec_into(self, x: torch.Tensor, out: torch.Tensor, sync=True):
        self.ctx.check(self.lib.qlg_matvec(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr())))
        if sync:
            self.ctx.synchronize()

    def __matmul__(self, state: BlockTensor) -> BlockTensor:
        x = self.adopt(state)
        out = self.new_vector(zero=False)
        torch.cuda.synchronize()
        self.matvec_into(x.data, out.data)
        return out

    def lanczos(self, init_state: BlockTensor, params: LanczosParams) -> LanczosRes:
        v = self.adopt(init_state)
        work = self.new_vector(zero=False)
        work.data.copy_(v.data)
        torch.cuda.synchronize()
        res = L.LanczosResult()
        prm = L.LanczosParams(params.error, params.max_iterations)
        self.ctx.check(self.lib.qlg_lanczos_gs(self.h, C.c_void_p(work.data.data_ptr()), C.byref(prm), C.byref(res)))
        return LanczosRes(int(res.iters), float(res.gs_eng), work)

    def expmv(self, init_state: BlockTensor, step_length: float, params: LanczosParams) -> ExpmvRes:
        v = self.adopt(init_state)
        work = self.new_vector(zero=False)
        work.data.copy_(v.data)
        torch.cuda.synchronize()
        it = C.c_int64()
        prm = L.LanczosParams(params.error, params.max_iterations)
        self.ctx.check(self.lib.qlg_lanczos_expmv(self.h, C.c_void_p(work.data.data_ptr()), float(step_length),
                                                  C.byref(prm), C.byref(it)))
        return ExpmvRes(int(it.value), work)

    def close(self):
        if getattr(self, "h", None):
            self.lib.qlg_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedEffectiveHamiltonian:
    """Independent chains' effective Hamiltonians solved together
    (``qlg_plan_batch``): each matvec stage of all members is one grouped
    launch, and ``lanczos`` runs every member's Lanczos solve in lock step
    (iteration counts and energies are those of separate solves).  With
    ``graphs=True`` each Lanczos iteration is captured as a hipGraph on first
    use and replayed by later solves on this batch."""

    def __init__(self, members: Sequence[EffectiveHamiltonian], graphs: bool = False):
        self.members = list(members)
        self.ctx = self.members[0].ctx
        self.lib = self.ctx.lib
        n = len(self.members)
        plans = (C.c_void_p * n)(*[m.h.value if isinstance(m.h, C.c_void_p) else m.h for m in self.members])
        h = C.c_void_p()
        torch.cuda.synchronize()
        self.ctx.check(self.lib.qlg_plan_batch(self.ctx.h, n, plans, C.byref(h)))
        self.h = h
        cnt = C.c_int64()
        offs = (C.c_int64 * (n + 1))()
        self.ctx.check(self.lib.qlg_plan_batch_offsets(h, C.byref(cnt), offs))
        self.offsets = list(offs)
        self.size = self.offsets[-1]
        self.device = self.members[0].device
        self.dtype = self.members[0].dtype
        if graphs:
            self.ctx.check(self.lib.qlg_plan_use_graphs(h, 1))

    def pack(self, states: Sequence[BlockTensor]) -> torch.Tensor:
        buf = torch.zeros(max(self.size, 1), dtype=self.dtype, device=self.device)
        for m, st, o in zip(self.members, states, self.offsets):
            v = m.adopt(st)
            buf[o:o + m.size] = v.data[:m.size]
        return buf

    def unpack(self, buf: torch.Tensor) -> List[BlockTensor]:
        out = []
        for m, o in zip(self.members, self.offsets):
            out.append(BlockTensor(m.axes, m.layout_coords, buf[o:o + max(m.size, 1)].clone(), m.layout_offsets))
        return out

    def flops(self):
        a, e = C.c_double(), C.c_double()
        self.ctx.check(self.lib.qlg_plan_flops(self.h, C.byref(a), C.byref(e)))
        return a.value, e.value

    def enable_timing(self, on=True):
        self.ctx.check(self.lib.qlg_plan_enable_timing(self.h, int(on)))

    def stage_ms(self):
        out = (C.c_double * 3)()
        self.ctx.check(self.lib.qlg_plan_stage_ms(self.h, out))
        return list(out)

    def matvec_into(self, x: torch.Tensor, out: torch.Tensor, sync=True):
        self.ctx.check(self.lib.qlg_matvec(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr())))
        if sync:
            self.ctx.synchronize()

    def lanczos(self, init_states: Sequence[BlockTensor], params: LanczosParams) -> List[LanczosRes]:
        work = self.pack(init_states)
        torch.cuda.synchronize()
        n = len(self.members)
        res = (L.LanczosResult * n)()
        prm = L.LanczosParams(params.error, params.max_iterations)
        self.ctx.check(self.lib.qlg_lanczos_gs_batch(self.h, C.c_void_p(work.data_ptr()), C.byref(prm), res))
        return [LanczosRes(int(r.iters), float(r.gs_eng), v) for r, v in zip(res, self.unpack(work))]

    def close(self):
        if getattr(self, "h", None):
            self.lib.qlg_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# reference-named entry points
# ---------------------------------------------------------------------------
def eff_ham_mul_two_site_state(eff_ham: Sequence[BlockTensor], state: BlockTensor) -> BlockTensor:
    """``lanczos_vmps_solver_impl.h:260-289`` (plans, applies once, returns a new tensor)."""
    return EffectiveHamiltonian(eff_ham, state) @ state


def eff_ham_mul_single_site_state(eff_ham: Sequence[BlockTensor], state: BlockTensor) -> BlockTensor:
    """``lanczos_vmps_solver_impl.h:291-304``."""
    return EffectiveHamiltonian(eff_ham, state) @ state


def LanczosSolver(eff_ham, init_state: BlockTensor, matvec=None, params: LanczosParams = LanczosParams(),
                  precision: str = "fp64") -> LanczosRes:
    """``lanczos_vmps_solver_impl.h:114-248``; ``eff_ham`` may be an EffectiveHamiltonian.

    The DMRG overload ``LanczosSolver(SuperBlockHamiltonianTerms&, TenT*, const
    LanczosParams&)`` (``lanczos_dmrg_solver_impl.h:61-196``) is the call with
    the term list as ``eff_ham`` and the parameters in the third position."""
    if isinstance(matvec, LanczosParams):   # DMRG form: (terms, state, params)
        params, matvec = matvec, None
    H = eff_ham if isinstance(eff_ham, EffectiveHamiltonian) else EffectiveHamiltonian(eff_ham, init_state,
                                                                                       precision)
    return H.lanczos(init_state, params)


def LanczosExpmvSolver(eff_ham, init_state: BlockTensor, matvec=None, step_length: float = 0.0,
                       params: LanczosParams = LanczosParams()) -> ExpmvRes:
    """``lanczos_expmv_solver_impl.h:70-184``."""
    H = eff_ham if isinstance(eff_ham, EffectiveHamiltonian) else EffectiveHamiltonian(eff_ham, init_state)
    return H.expmv(init_state, step_length, params)


def TridiagGsSolver(a: Sequence[float], b: Sequence[float], n: int, jobz: str = "N"):
    """``lanczos_vmps_solver_impl.h:306-355`` through the library's FP64 QL solver."""
    lib = L.lib()
    A = (C.c_double * n)(*a[:n])
    B = (C.c_double * max(n - 1, 1))(*(list(b[:n - 1]) or [0.0]))
    e = C.c_double()
    v = (C.c_double * n)()
    L.check(lib.qlg_tridiag_gs(A, B, n, jobz.encode(), C.byref(e), v if jobz == "V" else None))
    return e.value, (np.array(v[:]) if jobz == "V" else None)


def TridiagExpme1Solver(a: Sequence[float], b: Sequence[float], n: int, step_length: float) -> np.ndarray:
    """``lanczos_expmv_solver_impl.h:199-240``."""
    lib = L.lib()
    A = (C.c_double * n)(*a[:n])
    B = (C.c_double * max(n - 1, 1))(*(list(b[:n - 1]) or [0.0]))
    out = (C.c_double * (2 * n))()
    L.check(lib.qlg_tridiag_expme1(A, B, n, float(step_length), out))
    z = np.array(out[:])
    return z[0::2] + 1j * z[1::2]
