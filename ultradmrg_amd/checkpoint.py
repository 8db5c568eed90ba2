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
renamed, so an interrupted dump never leaves a truncated file in place; a
whole-MPS dump is written into a fresh directory and swapped in as a unit
(``save_mps``).
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


def _fsync_dir(path: str):
    fd = os.open(path, os.O_RDONLY)
    try:
        os.fsync(fd)
    finally:
        os.close(fd)


def save_mps(mps_path: str, mps: List[BlockTensor], meta: Optional[dict] = None):
    """Dump every site tensor and a manifest as one unit.

    The dump is written into a fresh sibling directory ``<mps_path>.partial``;
    the manifest goes last and records a generation id, the number of sites
    and each file's size and body CRC32 (the file's own trailer).  Only then is the finished directory
    swapped in: the previous dump is renamed to ``<mps_path>.prev``, the new
    one to ``<mps_path>``, and ``.prev`` is removed.  A process killed at any
    point leaves either the previous complete dump or the new one readable
    (``load_mps`` falls back to ``.prev`` when the swap was cut between its two
    renames); a half-written dump is never read."""
    import shutil
    import uuid
    mps_path = os.path.abspath(mps_path)
    parent = os.path.dirname(mps_path)
    os.makedirs(parent, exist_ok=True)
    part, prev = mps_path + ".partial", mps_path + ".prev"
    if os.path.exists(part):
        shutil.rmtree(part)
    os.makedirs(part)
    files = []
    for i, t in enumerate(mps):
        payload = _pack_tensor(t)
        _atomic_write(mps_tensor_name(part, i), payload)
        # the file's own trailer: CRC32 of its body (checked again when it is unpacked)
        files.append({"name": os.path.basename(mps_tensor_name(part, i)), "bytes": len(payload),
                      "crc32": struct.unpack("<I", payload[-4:])[0]})
    man = {"format": "qlgt", "version": VERSION, "length": len(mps), "generation": uuid.uuid4().hex,
           "files": files}
    man.update(meta or {})
    _atomic_write(os.path.join(part, "manifest.json"), json.dumps(man, indent=1).encode())
    _fsync_dir(part)
    if os.path.exists(prev):
        shutil.rmtree(prev)
    if os.path.exists(mps_path):
        os.rename(mps_path, prev)
    os.rename(part, mps_path)
    _fsync_dir(parent)
    if os.path.exists(prev):
        shutil.rmtree(prev)


def _check_bonds(mps: List[BlockTensor]):
    """Neighbouring site tensors must share their bond: the right bond of site i
    is the inverse of the left bond of site i+1 (same sectors and degeneracies)."""
    for i in range(len(mps) - 1):
        a, b = mps[i].axes[-1], mps[i + 1].axes[0]
        if a.qns != b.qns or a.degs != b.degs or a.dir != -b.dir:
            raise CheckpointError(f"MPS checkpoint: bond between sites {i} and {i + 1} does not match")


def load_mps(mps_path: str, device="cuda") -> Tuple[List[BlockTensor], dict]:
    """Read a dump written by ``save_mps``: every file's size and body CRC32
    must match the manifest (and the body its CRC), and neighbouring sites
    must share their bonds."""
    mps_path = os.path.abspath(mps_path)
    man_path = os.path.join(mps_path, "manifest.json")
    if not os.path.exists(man_path):
        prev = mps_path + ".prev"
        if os.path.exists(os.path.join(prev, "manifest.json")):
            mps_path, man_path = prev, os.path.join(prev, "manifest.json")
        else:
            raise CheckpointError(f"no complete MPS checkpoint in {mps_path}")
    with open(man_path) as f:
        man = json.load(f)
    n = int(man["length"])
    files = man.get("files")
    if files is None or len(files) != n:
        raise CheckpointError("MPS checkpoint manifest does not list every site file")
    mps = []
    for i in range(n):
        path = mps_tensor_name(mps_path, i)
        if not os.path.exists(path):
            raise CheckpointError(f"MPS checkpoint: site file {i} is missing")
        with open(path, "rb") as f:
            buf = f.read()
        if len(buf) != int(files[i]["bytes"]) or struct.unpack("<I", buf[-4:])[0] != int(files[i]["crc32"]):
            raise CheckpointError(f"MPS checkpoint: site file {i} does not belong to this dump "
                                  f"(generation {man.get('generation')})")
        mps.append(_unpack_tensor(buf, device))
    _check_bonds(mps)
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
