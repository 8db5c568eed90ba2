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


def _scaled(mps, c):
    from ultradmrg_amd.tensor import BlockTensor
    return [BlockTensor(t.axes, t.coords, t.data * c, t.offsets) for t in mps]


def test_interrupted_second_dump_keeps_the_first(tmp_path, monkeypatch):
    """A second dump killed part-way (here: after two site files) leaves the first
    dump readable and unchanged; nothing of the partial dump is read."""
    mps = [to_device(t, device="cpu") for t in _mps("heisenberg")]
    d = str(tmp_path / "mps")
    ck.save_mps(d, mps, {"sweeps_done": 1})
    real = ck._pack_tensor
    calls = [0]

    def dying_pack(t):
        calls[0] += 1
        if calls[0] == 3:
            raise KeyboardInterrupt("killed during the dump")
        return real(t)

    monkeypatch.setattr(ck, "_pack_tensor", dying_pack)
    with pytest.raises(KeyboardInterrupt):
        ck.save_mps(d, _scaled(mps, 2.0), {"sweeps_done": 2})
    monkeypatch.setattr(ck, "_pack_tensor", real)
    back, man = ck.load_mps(d, "cpu")
    assert man["sweeps_done"] == 1
    for a, b in zip(back, mps):
        _same(a, b)
    # a later complete dump replaces it and clears the partial directory
    ck.save_mps(d, _scaled(mps, 2.0), {"sweeps_done": 2})
    assert not os.path.exists(d + ".partial") and not os.path.exists(d + ".prev")
    back, man = ck.load_mps(d, "cpu")
    assert man["sweeps_done"] == 2
    assert torch.equal(back[0].data, mps[0].data * 2.0)


def test_swap_cut_between_renames_reads_previous(tmp_path):
    """If the swap is cut after the old dump was renamed to .prev and before the
    new one took its place, the loader reads the previous complete dump."""
    mps = [to_device(t, device="cpu") for t in _mps("hubbard")]
    d = str(tmp_path / "mps")
    ck.save_mps(d, mps, {"sweeps_done": 4})
    os.rename(d, d + ".prev")
    back, man = ck.load_mps(d, "cpu")
    assert man["sweeps_done"] == 4
    for a, b in zip(back, mps):
        _same(a, b)


def test_mixed_generation_and_bond_mismatch_rejected(tmp_path):
    """A site file from another dump (size/CRC not in the manifest) and a bond
    that does not match its neighbour are both refused."""
    mps = [to_device(t, device="cpu") for t in _mps("heisenberg")]
    d1, d2 = str(tmp_path / "a"), str(tmp_path / "b")
    ck.save_mps(d1, mps)
    ck.save_mps(d2, _scaled(mps, 3.0))
    import shutil
    shutil.copy(ck.mps_tensor_name(d2, 2), ck.mps_tensor_name(d1, 2))
    with pytest.raises(ck.CheckpointError, match="does not belong"):
        ck.load_mps(d1, "cpu")
    bad = list(mps)
    bad[3] = to_device(o_random_mps(o_heis(8), (0,), 4, seed=5)[3], device="cpu")
    ck.save_mps(d2, bad)
    with pytest.raises(ck.CheckpointError, match="bond"):
        ck.load_mps(d2, "cpu")
