"""CPU baseline: ctypes face of ``oracle/_build/libqlcpu.so`` (``heff_cpu.cpp``),
the reference's four-step two-site matvec
(``include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:260-289``) as
block-sparse contractions on the host with OpenMP and a dlopen'd BLAS.

TEST AND BASELINE INFRASTRUCTURE ONLY: tests/ check it against the numpy
oracle and bench.py times it as the CPU baseline; the product never loads it.
Built by ``make -C oracle`` (``__graft_entry__.build()`` runs it).
"""
from __future__ import annotations

import ctypes as C
import os
import platform
import subprocess
from typing import Optional, Sequence

import numpy as np

from .bsten import BSTensor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libqlcpu.so")

_lib = None
_blas = None


class _Desc(C.Structure):
    """The qlg_bsten layout (include/qlgpu.h), host data pointer."""
    _fields_ = [("rank", C.c_int32), ("qn_len", C.c_int32), ("nsect", C.POINTER(C.c_int32)),
                ("sect_qn", C.POINTER(C.c_int32)), ("sect_deg", C.POINTER(C.c_int64)),
                ("dir", C.POINTER(C.c_int8)), ("nblocks", C.c_int64), ("blk_coords", C.POINTER(C.c_int32)),
                ("blk_offset", C.POINTER(C.c_int64)), ("dtype", C.c_int32), ("data", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.qlcpu_load_blas.restype = C.c_int
        L.qlcpu_load_blas.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        L.qlcpu_threads.restype = C.c_int
        L.qlcpu_plan_two_site.restype = C.c_void_p
        L.qlcpu_plan_two_site.argtypes = [C.POINTER(_Desc)] * 5 + [C.POINTER(C.c_double)]
        L.qlcpu_out_layout.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.qlcpu_matvec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.qlcpu_plan_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _bundled_openblas() -> Optional[str]:
    """The OpenBLAS that scipy's wheel ships (symbols prefixed scipy_), if any."""
    try:
        import scipy
    except ImportError:
        return None
    d = os.path.realpath(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs"))
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if f.startswith("libscipy_openblas") and "64_" not in f and f.endswith(".so"):
            return os.path.join(d, f)
    return None


def load_blas() -> str:
    """Load a host BLAS once; returns its description."""
    global _blas
    if _blas is None:
        buf = C.create_string_buffer(512)
        path = _bundled_openblas() or ""
        lib().qlcpu_load_blas(path.encode(), b"scipy_", buf, 512)
        _blas = buf.value.decode()
    return _blas


def host_info() -> dict:
    model = platform.processor() or "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "omp_threads": lib().qlcpu_threads(), "blas": load_blas()}


class _Keep:
    def __init__(self, t: BSTensor):
        keys = sorted(t.blocks)
        r = t.rank
        self.keys = keys
        self.nsect = (C.c_int32 * r)(*[ix.nsect for ix in t.indexes])
        qn = [v for ix in t.indexes for s in range(ix.nsect) for v in ix.qn(s)]
        self.qn_len = len(t.indexes[0].qn(0))
        self.sect_qn = (C.c_int32 * max(len(qn), 1))(*qn)
        degs = [ix.deg(s) for ix in t.indexes for s in range(ix.nsect)]
        self.sect_deg = (C.c_int64 * max(len(degs), 1))(*degs)
        self.dirs = (C.c_int8 * r)(*[ix.dir for ix in t.indexes])
        flat = [v for k in keys for v in k]
        self.coords = (C.c_int32 * max(len(flat), 1))(*flat)
        sizes = [t.blocks[k].size for k in keys]
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if sizes else np.zeros(0, np.int64)
        self.offs = (C.c_int64 * max(len(offs), 1))(*offs.tolist())
        self.data = (np.ascontiguousarray(np.concatenate([t.blocks[k].reshape(-1) for k in keys]))
                     if keys else np.zeros(1)).astype(np.float64)
        self.desc = _Desc(r, self.qn_len, self.nsect, self.sect_qn, self.sect_deg, self.dirs, len(keys),
                          self.coords, self.offs, 0, self.data.ctypes.data_as(C.c_void_p))


class CpuTwoSite:
    """The CPU baseline's plan for [L, W1, W2, R] and states shaped like psi
    (real FP64)."""

    def __init__(self, eff: Sequence[BSTensor], psi: BSTensor):
        if any(t.dtype.kind == "c" for t in list(eff) + [psi]):
            raise ValueError("the CPU baseline is real FP64 only")
        load_blas()
        self._ops = [_Keep(t) for t in eff]
        self._psi = _Keep(psi)
        fl = C.c_double()
        self.h = lib().qlcpu_plan_two_site(*[C.byref(k.desc) for k in self._ops], C.byref(self._psi.desc),
                                           C.byref(fl))
        if not self.h:
            raise RuntimeError("qlcpu_plan_two_site failed")
        self.flops = fl.value
        nb, size = C.c_int64(), C.c_int64()
        lib().qlcpu_out_layout(self.h, C.byref(nb), C.byref(size), None, None)
        coords = (C.c_int32 * max(nb.value * 4, 1))()
        offs = (C.c_int64 * max(nb.value, 1))()
        lib().qlcpu_out_layout(self.h, None, None, coords, offs)
        self.out_keys = [tuple(coords[4 * b:4 * b + 4]) for b in range(nb.value)]
        self.out_offs = list(offs[:nb.value])
        self.out_size = size.value
        self.out_indexes = [eff[0].indexes[0], eff[1].indexes[2], eff[2].indexes[2], eff[3].indexes[0]]
        self.psi_data = self._psi.data
        self.out = np.zeros(max(self.out_size, 1))

    def matvec(self, psi_data: Optional[np.ndarray] = None) -> np.ndarray:
        x = self.psi_data if psi_data is None else np.ascontiguousarray(psi_data, dtype=np.float64)
        lib().qlcpu_matvec(self.h, x.ctypes.data_as(C.c_void_p), self.out.ctypes.data_as(C.c_void_p))
        return self.out

    def result(self) -> BSTensor:
        t = BSTensor(self.out_indexes, np.float64)
        for k, o in zip(self.out_keys, self.out_offs):
            shape = tuple(ix.deg(s) for ix, s in zip(self.out_indexes, k))
            t.blocks[k] = self.out[o:o + int(np.prod(shape))].reshape(shape).copy()
        return t

    def __del__(self):
        try:
            if self.h:
                lib().qlcpu_plan_destroy(self.h)
        except Exception:
            pass
