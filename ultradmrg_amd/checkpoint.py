"""Checkpoint / resume of GPU-resident MPS, and signal-driven stops.

The reference dumps each MPS tensor with qlten's binary stream operators
(``include/qlmps/one_dim_tn/framework/ten_vec.h:61-105``, file names
``mps_ten<i>`` from ``include/qlmps/consts.h:37``) and lets a running DMRG
stop on signals (``include/qlmps/algorithm/dmrg/dmrg_impl.h:478-502``: SIGUSR1
/ SIGUSR2 stop after the current sweep, SIGINT / SIGTERM stop at once).  The
qlten format itself is not part of the reference, so this module defines its
own, versioned format for a block-sparse tensor:

    magic   b"QLGT" | version u32 | rank u32 | qn_len u32 | dtype u32 (0 f64, 1 c128)
    per axis: dir i32 | nsect u32 | qn i32[nsect * qn_len] | deg i64[nsect]
    nblocks u64 | coords i32[nblocks * rank] | offsets i64[nblocks]
    nelem u64 | data (little-endian f64 or interleaved c128)
    crc32 u32 of everything before it

All integers are little-endian.  Files are written to a temporary name and
renamed, so an interrupted dump never leaves a truncated file in place.
"""
from __future__ import annotations

import json
import os
import signal
import struct
import zlib
from typing import List, Optional, Tuple

import numpy as np
import torch

from .tensor import Axis, BlockTensor

MAGIC = b"QLGT"
VERSION = 1


class CheckpointError(RuntimeError):
    pass


def _pack_tensor(t: BlockTensor) -> bytes:
    cplx = t.data.is_complex()
    qn_len = len(t.axes[0].qns[0])
    out = [MAGIC, struct.pack("<IIII", VERSION, t.rank, qn_len, 1 if cplx else 0)]
    for ax in t.axes:
        out.append(struct.pack("<iI", ax.dir, ax.nsect))
        out.append(np.asarray(ax.qns, dtype="<i4").reshape(-1).tobytes())
        out.append(np.asarray(ax.degs, dtype="<i8").tobytes())
    out.append(struct.pack("<Q", len(t.coords)))
    out.append(np.asarray(t.coords, dtype="<i4").reshape(-1).tobytes())
    out.append(np.asarray(t.offsets, dtype="<i8").tobytes())
    data = t.data.detach().cpu().numpy()
    data = data.astype("<c16" if cplx else "<f8", copy=False)
    out.append(struct.pack("<Q", data.size))
    out.append(data.tobytes())
    body = b"".join(out)
    return body + struct.pack("<I", zlib.crc32(body) & 0xFFFFFFFF)


def _unpack_tensor(buf: bytes, device) -> BlockTensor:
    if len(buf) < 24 or buf[:4] != MAGIC:
        raise CheckpointError("not a tensor checkpoint (bad magic)")
    body, (crc,) = buf[:-4], struct.unpack("<I", buf[-4:])
    if zlib.crc32(body) & 0xFFFFFFFF != crc:
        raise CheckpointError("tensor checkpoint is corrupt (CRC mismatch)")
    pos = 4
    version, rank, qn_len, dtype = struct.unpack_from("<IIII", body, pos)
    pos += 16
    if version != VERSION:
        raise CheckpointError(f"unsupported checkpoint version {version}")
    axes = []
    for _ in range(rank):
        d, ns = struct.unpack_from("<iI", body, pos)
        pos += 8
        qn = np.frombuffer(body, dtype="<i4", count=ns * qn_len, offset=pos).reshape(ns, qn_len)
        pos += 4 * ns * qn_len
        deg = np.frombuffer(body, dtype="<i8", count=ns, offset=pos)
        pos += 8 * ns
        axes.append(Axis(tuple(tuple(int(v) for v in q) for q in qn), tuple(int(v) for v in deg), int(d)))
    (nb,) = struct.unpack_from("<Q", body, pos)
    pos += 8
    coords = np.frombuffer(body, dtype="<i4", count=nb * rank, offset=pos).reshape(nb, rank)
    pos += 4 * nb * rank
    offs = np.frombuffer(body, dtype="<i8", count=nb, offset=pos)
    pos += 8 * nb
    (ne,) = struct.unpack_from("<Q", body, pos)
    pos += 8
    arr = np.frombuffer(body, dtype="<c16" if dtype == 1 else "<f8", count=ne, offset=pos)
    if pos + arr.nbytes != len(body):
        raise CheckpointError("tensor checkpoint has trailing or missing bytes")
    data = torch.from_numpy(arr.copy()).to(device)
    return BlockTensor(axes, [tuple(int(v) for v in c) for c in coords], data, [int(v) for v in offs])


def _atomic_write(path: str, payload: bytes):
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        f.write(payload)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def save_tensor(path: str, t: BlockTensor):
    _atomic_write(path, _pack_tensor(t))


def load_tensor(path: str, device="cuda") -> BlockTensor:
    with open(path, "rb") as f:
        return _unpack_tensor(f.read(), device)


def mps_tensor_name(mps_path: str, i: int) -> str:
    """``GenMPSTenName`` with this format's suffix: <mps_path>/mps_ten<i>.qlgt."""
    return os.path.join(mps_path, f"mps_ten{i}.qlgt")


def save_mps(mps_path: str, mps: List[BlockTensor], meta: Optional[dict] = None):
    """Dump every site tensor, then the manifest (written last: a directory with
    a manifest always holds a complete MPS)."""
    os.makedirs(mps_path, exist_ok=True)
    for i, t in enumerate(mps):
        save_tensor(mps_tensor_name(mps_path, i), t)
    man = {"format": "qlgt", "version": VERSION, "length": len(mps)}
    man.update(meta or {})
    _atomic_write(os.path.join(mps_path, "manifest.json"), json.dumps(man, indent=1).encode())


def load_mps(mps_path: str, device="cuda") -> Tuple[List[BlockTensor], dict]:
    man_path = os.path.join(mps_path, "manifest.json")
    if not os.path.exists(man_path):
        raise CheckpointError(f"no complete MPS checkpoint in {mps_path}")
    with open(man_path) as f:
        man = json.load(f)
    mps = [load_tensor(mps_tensor_name(mps_path, i), device) for i in range(int(man["length"]))]
    return mps, man


# ---------------------------------------------------------------------------
# signal-driven stop
# ---------------------------------------------------------------------------
class StopFlags:
    """SIGUSR1 / SIGUSR2: stop after the current sweep; SIGINT / SIGTERM: stop
    after the current update (the driver checkpoints before returning)."""

    def __init__(self):
        self.stop = False
        self.emergency = False
        self._old = {}

    def _handler(self, sig, frame):
        self.stop = True
        if sig in (signal.SIGINT, signal.SIGTERM):
            self.emergency = True

    def install(self):
        for s in (signal.SIGUSR1, signal.SIGUSR2, signal.SIGINT, signal.SIGTERM):
            self._old[s] = signal.signal(s, self._handler)
        return self

    def uninstall(self):
        for s, h in self._old.items():
            signal.signal(s, h)
        self._old.clear()
