"""Device-resident block-sparse tensors and the Python face of the C ABI.

A ``BlockTensor`` holds the quantum-number bookkeeping on the host (axes with
ordered sectors and a direction, a block list with element offsets) and one
flat ``torch`` buffer on the GPU.  Blocks are dense row-major; the canonical
order is lexicographic in the sector coordinates with offsets equal to the
prefix sums of block sizes (the layout the planner uses for Krylov vectors).
The same storage model is what the reference's QLTensor keeps, so the
descriptor handed to the library is zero-copy.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L

IN, OUT = -1, +1


@dataclass(frozen=True)
class Axis:
    qns: Tuple[Tuple[int, ...], ...]
    degs: Tuple[int, ...]
    dir: int

    @staticmethod
    def make(sectors: Sequence[Tuple[Sequence[int], int]], dir: int) -> "Axis":
        return Axis(tuple(tuple(int(v) for v in q) for q, _ in sectors), tuple(int(d) for _, d in sectors), int(dir))

    def inverse(self) -> "Axis":
        return Axis(self.qns, self.degs, -self.dir)

    @property
    def nsect(self):
        return len(self.qns)

    @property
    def dim(self):
        return sum(self.degs)


def _flow(axes: Sequence[Axis], coords) -> Tuple[int, ...]:
    q = None
    for ax, s in zip(axes, coords):
        t = ax.qns[s] if ax.dir == OUT else tuple(-v for v in ax.qns[s])
        q = t if q is None else tuple(a + b for a, b in zip(q, t))
    return q


class BlockTensor:
    """Block-sparse tensor: host bookkeeping + one flat device buffer."""

    def __init__(self, axes: Sequence[Axis], coords: Sequence[Tuple[int, ...]], data: torch.Tensor,
                 offsets: Optional[Sequence[int]] = None):
        self.axes = list(axes)
        self.coords = [tuple(int(v) for v in c) for c in coords]
        sizes = [self.block_size(c) for c in self.coords]
        if offsets is None:
            offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if sizes else []
        self.offsets = [int(o) for o in offsets]
        self.size = int(max([o + s for o, s in zip(self.offsets, sizes)], default=0))
        if data.numel() < self.size:
            raise ValueError("data buffer smaller than the block layout")
        self.data = data
        self._desc_keep = None

    # ----- structure ------------------------------------------------------------
    @property
    def rank(self):
        return len(self.axes)

    @property
    def qn_len(self):
        return len(self.axes[0].qns[0])

    @property
    def is_complex(self):
        return self.data.is_complex()

    def block_shape(self, c):
        return tuple(ax.degs[s] for ax, s in zip(self.axes, c))

    def block_size(self, c):
        n = 1
        for ax, s in zip(self.axes, c):
            n *= ax.degs[s]
        return n

    # ----- construction -------------------------------------------------------------
    @classmethod
    def from_blocks(cls, axes: Sequence[Axis], blocks: Dict[Tuple[int, ...], np.ndarray], device="cuda",
                    dtype=None) -> "BlockTensor":
        keys = sorted(blocks)
        parts = [np.asarray(blocks[k]).reshape(-1) for k in keys]
        flat = np.concatenate(parts) if parts else np.zeros(0)
        if dtype is None:
            dtype = torch.complex128 if np.iscomplexobj(flat) else torch.float64
        t = torch.from_numpy(np.ascontiguousarray(flat)).to(dtype=dtype, device=device)
        return cls(axes, keys, t)

    @classmethod
    def zeros_like_layout(cls, axes, coords, offsets, size, dtype, device="cuda") -> "BlockTensor":
        return cls(axes, coords, torch.zeros(max(size, 1), dtype=dtype, device=device), offsets)

    def blocks(self) -> Dict[Tuple[int, ...], np.ndarray]:
        host = self.data.detach().cpu().numpy()
        return {c: host[o:o + self.block_size(c)].reshape(self.block_shape(c))
                for c, o in zip(self.coords, self.offsets)}

    # ----- C descriptor ------------------------------------------------------------
    def descriptor(self) -> L.Bsten:
        r = self.rank
        nsect = (C.c_int32 * r)(*[ax.nsect for ax in self.axes])
        qn = [v for ax in self.axes for q in ax.qns for v in q]
        sect_qn = (C.c_int32 * max(len(qn), 1))(*qn)
        degs = [d for ax in self.axes for d in ax.degs]
        sect_deg = (C.c_int64 * max(len(degs), 1))(*degs)
        dirs = (C.c_int8 * r)(*[ax.dir for ax in self.axes])
        flat = [v for c in self.coords for v in c]
        coords = (C.c_int32 * max(len(flat), 1))(*flat)
        offs = (C.c_int64 * max(len(self.offsets), 1))(*self.offsets)
        d = L.Bsten(r, self.qn_len, nsect, sect_qn, sect_deg, dirs, len(self.coords), coords, offs,
                    L.QLG_C128 if self.is_complex else L.QLG_F64, C.c_void_p(self.data.data_ptr()))
        self._desc_keep = (nsect, sect_qn, sect_deg, dirs, coords, offs, d)
        return d


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
class Context:
    """One library context (device + HIP stream)."""

    def __init__(self, device: int = 0):
        self.lib = L.lib()
        h = C.c_void_p()
        L.check(self.lib.qlg_ctx_create(device, None, C.byref(h)))
        self.h = h
        self.device = device

    def check(self, rc):
        L.check(rc, self.h)

    def synchronize(self):
        self.check(self.lib.qlg_ctx_synchronize(self.h))

    def memory_stats(self) -> dict:
        """The device block cache: cached bytes and its hipMalloc / reuse / hipFree counts."""
        st = (C.c_int64 * 4)()
        self.check(self.lib.qlg_ctx_memory_stats(self.h, st))
        return {"cached_bytes": st[0], "mallocs": st[1], "reuses": st[2], "frees": st[3]}

    def release_cached(self):
        """Return the cached device blocks to the driver."""
        self.check(self.lib.qlg_ctx_release_cached(self.h))

    def close(self):
        if self.h:
            self.lib.qlg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP path has no CPU fallback")
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


def dtype_code(t: torch.Tensor) -> int:
    return L.QLG_C128 if t.is_complex() else L.QLG_F64


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def dot(x: torch.Tensor, y: torch.Tensor, ctx: Context = None):
    """sum x * conj(y) (deterministic)."""
    ctx = ctx or default_context(x.device.index or 0)
    out = (C.c_double * 2)()
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_dot(ctx.h, dtype_code(x), x.numel(), _ptr(x), _ptr(y), out))
    return complex(out[0], out[1]) if x.is_complex() else out[0]


def nrm2(x: torch.Tensor, ctx: Context = None) -> float:
    ctx = ctx or default_context(x.device.index or 0)
    out = (C.c_double * 1)()
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_nrm2(ctx.h, dtype_code(x), x.numel(), _ptr(x), out))
    return out[0]


def _scalar(v, cplx):
    v = complex(v)
    return (C.c_double * 2)(v.real, v.imag) if cplx else (C.c_double * 1)(v.real)


def scal(alpha, x: torch.Tensor, ctx: Context = None):
    ctx = ctx or default_context(x.device.index or 0)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_scal(ctx.h, dtype_code(x), x.numel(), _scalar(alpha, x.is_complex()), _ptr(x)))
    ctx.synchronize()


def axpy(alpha, x: torch.Tensor, y: torch.Tensor, ctx: Context = None):
    ctx = ctx or default_context(x.device.index or 0)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_axpy(ctx.h, dtype_code(x), x.numel(), _scalar(alpha, x.is_complex()), _ptr(x), _ptr(y)))
    ctx.synchronize()


def linear_combine(coefs, bases: Sequence[torch.Tensor], beta, y: torch.Tensor, ctx: Context = None):
    """y <- beta*y + sum_i coefs[i]*bases[i] (qlten LinearCombine)."""
    ctx = ctx or default_context(y.device.index or 0)
    cplx = y.is_complex()
    k = len(bases)
    if cplx:
        cf = (C.c_double * (2 * max(k, 1)))(*[v for c in coefs for v in (complex(c).real, complex(c).imag)])
    else:
        cf = (C.c_double * max(k, 1))(*[float(c) for c in coefs])
    ptrs = (C.c_void_p * max(k, 1))(*[b.data_ptr() for b in bases])
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_linear_combine(ctx.h, dtype_code(y), y.numel(), k, cf, ptrs, _scalar(beta, cplx), _ptr(y)))


# ---------------------------------------------------------------------------
# contractions around the Krylov solve
# ---------------------------------------------------------------------------
class _OpPlan:
    """A tensordot / environment-update plan and its result layout."""

    def __init__(self, ctx: Context, handle, axes, rank):
        self.ctx, self.h, self.axes = ctx, handle, axes
        lib = ctx.lib
        nb, size = C.c_int64(), C.c_int64()
        ctx.check(lib.qlg_plan_vector_layout(handle, C.byref(nb), C.byref(size), None, None))
        coords = (C.c_int32 * max(nb.value * rank, 1))()
        offs = (C.c_int64 * max(nb.value, 1))()
        ctx.check(lib.qlg_plan_vector_layout(handle, None, None, coords, offs))
        self.coords = [tuple(coords[b * rank:(b + 1) * rank]) for b in range(nb.value)]
        self.offsets = list(offs[:nb.value])
        self.size = size.value

    def run(self, in0: BlockTensor, in1: BlockTensor, dtype) -> BlockTensor:
        out = torch.empty(max(self.size, 1), dtype=dtype, device=in0.data.device)
        self.ctx.check(self.ctx.lib.qlg_execute(self.h, C.c_void_p(in0.data.data_ptr()),
                                                C.c_void_p(in1.data.data_ptr()), C.c_void_p(out.data_ptr())))
        self.ctx.synchronize()
        return BlockTensor(self.axes, self.coords, out, self.offsets)

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.qlg_plan_destroy(self.h)
        except Exception:
            pass


def contract(A: BlockTensor, B: BlockTensor, k: int = 1, conj_a=False, conj_b=False,
             ctx: Context = None) -> BlockTensor:
    """C = A x B over A's last k axes and B's first k axes (result: A free axes, then B's)."""
    ctx = ctx or default_context(A.data.device.index or 0)
    h = C.c_void_p()
    da, db = A.descriptor(), B.descriptor()
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_plan_contract(ctx.h, C.byref(da), C.byref(db), k, int(conj_a), int(conj_b), C.byref(h)))
    axes = A.axes[:A.rank - k] + B.axes[k:]
    plan = _OpPlan(ctx, h, axes, len(axes))
    dtype = torch.complex128 if (A.is_complex or B.is_complex) else torch.float64
    return plan.run(A, B, dtype)


def transpose(t: BlockTensor, perm: Sequence[int], ctx: Context = None) -> BlockTensor:
    """The tensor with axis i of the result = axis perm[i] of t (qlten's
    Transpose, as at two_site_update_finite_tdvp_impl.h:36), on the GPU."""
    ctx = ctx or default_context(t.data.device.index or 0)
    perm = [int(p) for p in perm]
    h = C.c_void_p()
    d = t.descriptor()
    pa = (C.c_int32 * len(perm))(*perm)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_plan_transpose(ctx.h, C.byref(d), pa, C.byref(h)))
    plan = _OpPlan(ctx, h, [t.axes[p] for p in perm], t.rank)
    return plan.run(t, t, t.data.dtype)


def env_update(env: BlockTensor, mps: BlockTensor, mpo: BlockTensor, side: str,
               ctx: Context = None) -> BlockTensor:
    """UpdateSiteLenvs (side 'l') / UpdateSiteRenvs (side 'r') on the GPU."""
    ctx = ctx or default_context(env.data.device.index or 0)
    h = C.c_void_p()
    de, dm, dw = env.descriptor(), mps.descriptor(), mpo.descriptor()
    torch.cuda.synchronize()
    ctx.check(ctx.lib.qlg_plan_env(ctx.h, C.byref(de), C.byref(dm), C.byref(dw), side.encode(), C.byref(h)))
    if side == "l":
        axes = [mps.axes[2].inverse(), mpo.axes[3], mps.axes[2]]
    else:
        axes = [mps.axes[0].inverse(), mpo.axes[0], mps.axes[0]]
    plan = _OpPlan(ctx, h, axes, 3)
    return plan.run(env, mps, env.data.dtype)
