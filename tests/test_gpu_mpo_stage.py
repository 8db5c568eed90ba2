"""The matvec's MPO stage (Y = C . X per charge channel) on the shapes its
kernel has to handle: every tile shape the stage can run (pinned with
QLG_TILES; the planner itself picks among 16x128, 32x128 and 64x64), and
channels with more than 256 input rows, whose row descriptors go through
LDS one 256-row window at a time.  Each matvec is checked against the oracle
at the precision's bar: FP64 1e-12 max-abs relative, the split modes their
Frobenius bars."""
import numpy as np
import pytest
import torch

from oracle.bsten import OUT, BSTensor, Index
from oracle.lanczos import eff_ham_mul_two_site_state
from oracle.models import build_mpo, heisenberg_chain, phys_index
from tests.convert import to_device, to_oracle
from tests.workloads import gaussian_bond, synthetic_eff_ham

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd.solver import EffectiveHamiltonian

TOL = {"fp64": 1e-12, "fp32": 2e-6, "bf16x6": 2e-6, "bf16x3": 2e-4}


def _rel(a, b, frob):
    assert sorted(a.blocks) == sorted(b.blocks), "block structure differs"
    if frob:
        num = sum(float(np.sum(np.abs(a.blocks[k] - b.blocks[k]) ** 2)) for k in a.blocks)
        den = sum(float(np.sum(np.abs(b.blocks[k]) ** 2)) for k in b.blocks)
        return (num / den) ** 0.5
    num = max(float(np.abs(a.blocks[k] - b.blocks[k]).max()) for k in a.blocks)
    den = max(float(np.abs(b.blocks[k]).max()) for k in b.blocks)
    return num / den


def _complexify(t, rng):
    c = BSTensor(t.indexes, np.complex128, t.div())
    c.blocks = {k: v + 1j * rng.uniform(-1, 1, size=v.shape) for k, v in t.blocks.items()}
    return c


def _check(eff, psi, prec):
    ref = eff_ham_mul_two_site_state(eff, psi)
    H = EffectiveHamiltonian([to_device(t) for t in eff], to_device(psi), precision=prec)
    out = H @ to_device(psi)
    return _rel(to_oracle(out, ref), ref, prec != "fp64")


def wide_bond_eff_ham(D, w, seed, nsect_half=2):
    """Spin-1/2 two-site problem whose MPO bonds carry one zero-charge sector of
    dimension w, with random site operators that conserve S^z.  A channel of
    total two-site charge 0 then has 2 w input rows (w >= 129 puts it past
    one 256-row descriptor window)."""
    rng = np.random.default_rng(seed)
    m = heisenberg_chain(8)
    W0 = build_mpo(m)[3]
    zero = (0,)
    wl = Index.make([(zero, w)], W0.indexes[0].dir)
    wm = Index.make([(zero, w)], W0.indexes[3].dir)
    wr = Index.make([(zero, w)], W0.indexes[3].dir)
    Wl = BSTensor([wl, W0.indexes[1], W0.indexes[2], wm], np.float64, div=zero)
    Wr = BSTensor([wm.inverse(), W0.indexes[1], W0.indexes[2], wr], np.float64, div=zero)
    for W in (Wl, Wr):
        for c in W.allowed_blocks(zero):
            W.blocks[c] = rng.uniform(-1, 1, size=W.blk_shape(c)) / w
    secs = gaussian_bond(m.qns, max(1.0, nsect_half / 2.5), D, nsect_half, 1)
    a_out = Index.make(secs, OUT)
    a_in, b_out = a_out.inverse(), Index.make(secs, OUT)
    b_in = b_out.inverse()
    pout = phys_index(m.qns, OUT)
    L = BSTensor([a_in, Wl.indexes[0].inverse(), a_out], np.float64, div=zero)
    R = BSTensor([b_out, Wr.indexes[3].inverse(), b_in], np.float64, div=zero)
    for t in (L, R):
        for c in t.allowed_blocks(zero):
            t.blocks[c] = rng.uniform(-1, 1, size=t.blk_shape(c))
    psi = BSTensor([a_in, pout, pout, b_out], np.float64, div=zero)
    for c in psi.allowed_blocks(zero):
        psi.blocks[c] = rng.uniform(-1, 1, size=psi.blk_shape(c))
    return [L, Wl, Wr, R], psi


@pytest.mark.parametrize("shape", ["16x128", "32x128", "64x128", "16x256", "32x256", "64x64"])
@pytest.mark.parametrize("prec", ["fp64", "bf16x6"])
def test_mpo_stage_every_tile_shape(monkeypatch, shape, prec):
    print(f"[case] MPO-stage tile {shape} {prec}", flush=True)
    monkeypatch.setenv("QLG_TILES", f"64x64,{shape},64x64")
    eff, psi = synthetic_eff_ham("hubbard", 160, seed=4, nsect_half=3)
    err = _check(eff, psi, prec)
    print(f"  error {err:.2e}", flush=True)
    assert err < TOL[prec]


@pytest.mark.parametrize("shape", ["16x128", "32x128", "64x64"])
def test_mpo_stage_every_tile_shape_complex(monkeypatch, shape):
    print(f"[case] MPO-stage tile {shape} complex", flush=True)
    monkeypatch.setenv("QLG_TILES", f"64x64,{shape},64x64")
    eff, psi = synthetic_eff_ham("hubbard", 96, seed=12, nsect_half=3)
    rng = np.random.default_rng(112)
    eff = [_complexify(t, rng) for t in eff]
    err = _check(eff, _complexify(psi, rng), "fp64")
    print(f"  error {err:.2e}", flush=True)
    assert err < TOL["fp64"]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x6", "bf16x3"])
@pytest.mark.parametrize("w,shape", [(150, "32x128"), (150, "16x256"), (300, "64x64")])
def test_mpo_stage_channels_past_descriptor_window(monkeypatch, capfd, prec, w, shape):
    """K_in = 300 and 600: two and three 256-row descriptor windows, the last
    one partial, and a last K chunk that is not a whole 32 rows."""
    print(f"[case] MPO-stage wide bond w={w} {shape} {prec}", flush=True)
    monkeypatch.setenv("QLG_TILES", f"64x64,{shape},64x64")
    monkeypatch.setenv("QLG_VERBOSE", "1")
    eff, psi = wide_bond_eff_ham(120 if w < 200 else 80, w, seed=7)
    err = _check(eff, psi, prec)
    log = capfd.readouterr().err
    with capfd.disabled():
        print(f"  error {err:.2e}", flush=True)
    line = [s for s in log.splitlines() if "MPO channels:" in s]
    assert line, "no MPO channel statistics printed"
    nbig = int(line[-1].split("kin>256:")[1].split(",")[0])
    assert nbig > 0, line[-1]
    assert err < TOL[prec]


def test_mpo_stage_channels_past_descriptor_window_complex(monkeypatch):
    print("[case] MPO-stage wide bond complex", flush=True)
    monkeypatch.setenv("QLG_TILES", "64x64,32x128,64x64")
    eff, psi = wide_bond_eff_ham(120, 150, seed=8)
    rng = np.random.default_rng(108)
    eff = [_complexify(t, rng) for t in eff]
    err = _check(eff, _complexify(psi, rng), "fp64")
    print(f"  error {err:.2e}", flush=True)
    assert err < TOL["fp64"]
