"""MPS checkpoint format and signal-driven stops, on the CPU.

The reference dumps site tensors one file per site (ten_vec.h:61-105, names
mps_ten<i> from consts.h:37) and stops a DMRG run on signals
(dmrg_impl.h:478-502).  Its on-disk format is qlten's, which is not part of the
reference sources, so these tests pin this package's own format: exact round trips
of real and complex tensors, detection of corruption and truncation, the
manifest written last, and the signal-to-flag mapping."""
import json
import os
import signal

import pytest
import torch

from oracle.dmrg import random_mps as o_random_mps
from oracle.models import heisenberg_chain as o_heis, hubbard_cylinder as o_hub
from tests.convert import to_device
from ultradmrg_amd import checkpoint as ck


def _mps(which):
    if which == "heisenberg":
        return o_random_mps(o_heis(8), (0,), 12, seed=5)
    return o_random_mps(o_hub(2, 2), (4, 0), 16, seed=6)


def _same(a, b):
    assert [ax.qns for ax in a.axes] == [ax.qns for ax in b.axes]
    assert [ax.degs for ax in a.axes] == [ax.degs for ax in b.axes]
    assert [ax.dir for ax in a.axes] == [ax.dir for ax in b.axes]
    assert list(a.coords) == list(b.coords) and list(a.offsets) == list(b.offsets)
    assert a.data.dtype == b.data.dtype
    assert torch.equal(a.data, b.data)


@pytest.mark.parametrize("which", ["heisenberg", "hubbard"])
@pytest.mark.parametrize("cplx", [False, True])
def test_tensor_roundtrip(tmp_path, which, cplx):
    for i, t in enumerate(_mps(which)):
        d = to_device(t, device="cpu")
        if cplx:
            g = torch.Generator().manual_seed(i)
            d.data = d.data.to(torch.complex128) + 1j * torch.rand(d.data.shape, generator=g, dtype=torch.float64)
        p = str(tmp_path / f"t{i}.qlgt")
        ck.save_tensor(p, d)
        _same(ck.load_tensor(p, "cpu"), d)


def test_corruption_detected(tmp_path):
    d = to_device(_mps("hubbard")[1], device="cpu")
    p = str(tmp_path / "t.qlgt")
    ck.save_tensor(p, d)
    raw = bytearray(open(p, "rb").read())
    flipped = bytearray(raw)
    flipped[len(raw) // 2] ^= 0x10
    open(p, "wb").write(bytes(flipped))
    with pytest.raises(ck.CheckpointError, match="CRC"):
        ck.load_tensor(p, "cpu")
    open(p, "wb").write(bytes(raw[:-9]))
    with pytest.raises(ck.CheckpointError):
        ck.load_tensor(p, "cpu")
    open(p, "wb").write(b"XXXX" + bytes(raw[4:]))
    with pytest.raises(ck.CheckpointError, match="magic"):
        ck.load_tensor(p, "cpu")


def test_mps_dir_roundtrip_and_manifest(tmp_path):
    mps = [to_device(t, device="cpu") for t in _mps("heisenberg")]
    d = str(tmp_path / "mps")
    with pytest.raises(ck.CheckpointError):
        ck.load_mps(d, "cpu")
    ck.save_mps(d, mps, {"sweeps_done": 3, "energy": -1.5, "complete_sweep": True})
    names = sorted(os.listdir(d))
    assert names == sorted(["manifest.json"] + [f"mps_ten{i}.qlgt" for i in range(len(mps))])
    man = json.load(open(os.path.join(d, "manifest.json")))
    assert man["length"] == len(mps) and man["sweeps_done"] == 3 and man["format"] == "qlgt"
    back, man2 = ck.load_mps(d, "cpu")
    assert man2 == man
    for a, b in zip(back, mps):
        _same(a, b)


def test_stop_flags():
    f = ck.StopFlags().install()
    try:
        os.kill(os.getpid(), signal.SIGUSR1)
        assert f.stop and not f.emergency
        f.stop = False
        os.kill(os.getpid(), signal.SIGUSR2)
        assert f.stop and not f.emergency
        os.kill(os.getpid(), signal.SIGTERM)
        assert f.stop and f.emergency
        f.stop = f.emergency = False
        os.kill(os.getpid(), signal.SIGINT)
        assert f.stop and f.emergency
    finally:
        f.uninstall()
    assert signal.getsignal(signal.SIGINT) is signal.default_int_handler
