"""Split-K stage 3 of the matvec: tasks with long K run as parts that write
partial panels, summed in part order by a reduce kernel.  Forced on small
cases with QLG_SPLITK (the chunk; every task longer than 3/2 of it is cut),
for every arithmetic variant that runs stage 3, against the oracle at the
variant's bar; the partial-panel path must give the same numbers as the
unsplit one to rounding (FP64: 1e-12 of the oracle; split modes: their
Frobenius bars)."""
import numpy as np
import pytest
import torch

from oracle.lanczos import eff_ham_mul_two_site_state
from tests.convert import to_device, to_oracle
from tests.workloads import synthetic_eff_ham

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


def _complexify(t, rng, real_valued=False):
    from oracle.bsten import BSTensor
    c = BSTensor(t.indexes, np.complex128, t.div())
    c.blocks = {k: (v + 0j if real_valued else v + 1j * rng.uniform(-1, 1, size=v.shape)) for k, v in t.blocks.items()}
    return c


@pytest.mark.parametrize("chunk", [64, 128])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x6", "bf16x3"])
@pytest.mark.parametrize("kind,D,nh,seed", [("heisenberg", 300, 6, 2), ("hubbard", 160, 3, 4)])
def test_split_k_matvec(monkeypatch, capfd, chunk, prec, kind, D, nh, seed):
    print(f"[case] split-K chunk {chunk} {prec} {kind} D={D}", flush=True)
    monkeypatch.setenv("QLG_SPLITK", str(chunk))
    monkeypatch.setenv("QLG_VERBOSE", "1")
    eff, psi = synthetic_eff_ham(kind, D, seed=seed, nsect_half=nh)
    ref = eff_ham_mul_two_site_state(eff, psi)
    H = EffectiveHamiltonian([to_device(t) for t in eff], to_device(psi), precision=prec)
    out = H @ to_device(psi)
    err = _rel(to_oracle(out, ref), ref, prec != "fp64")
    log = capfd.readouterr().err
    with capfd.disabled():
        print(f"  error {err:.2e}", flush=True)
    assert "stage 3 split-K: chunk" in log, "no stage-3 task was split"
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "bf16x6", "bf16x3"])
def test_split_k_complex_matvec(monkeypatch, capfd, prec):
    print(f"[case] split-K complex {prec}", flush=True)
    monkeypatch.setenv("QLG_SPLITK", "64")
    monkeypatch.setenv("QLG_VERBOSE", "1")
    eff, psi = synthetic_eff_ham("hubbard", 160, seed=4, nsect_half=3)
    rng = np.random.default_rng(104)
    L, Wl, Wr, R = eff
    eff = [_complexify(L, rng), _complexify(Wl, rng, True), _complexify(Wr, rng, True), _complexify(R, rng)]
    psi = _complexify(psi, rng)
    ref = eff_ham_mul_two_site_state(eff, psi)
    H = EffectiveHamiltonian([to_device(t) for t in eff], to_device(psi), precision=prec)
    out = H @ to_device(psi)
    err = _rel(to_oracle(out, ref), ref, prec != "fp64")
    log = capfd.readouterr().err
    with capfd.disabled():
        print(f"  error {err:.2e}", flush=True)
    assert "stage 3 split-K: chunk" in log
    assert err < TOL[prec]


def test_split_k_repeatable_and_tile_shapes(monkeypatch):
    """Every stage-3 tile shape with forced split-K matches the oracle, and
    two matvecs on the same plan are bit-identical (the partial panels are
    summed in a fixed order)."""
    eff, psi = synthetic_eff_ham("hubbard", 160, seed=4, nsect_half=3)
    ref = eff_ham_mul_two_site_state(eff, psi)
    monkeypatch.setenv("QLG_SPLITK", "128")
    for shape in ("32x32", "64x32", "32x64", "64x64", "128x64", "64x128", "128x128", "64x16", "128x16"):
        print(f"[case] split-K tile {shape}", flush=True)
        monkeypatch.setenv("QLG_TILES", f"64x64,32x128,{shape}")
        H = EffectiveHamiltonian([to_device(t) for t in eff], to_device(psi))
        d_psi = to_device(psi)
        a = H @ d_psi
        b = H @ d_psi
        assert torch.equal(a.data, b.data)
        assert _rel(to_oracle(a, ref), ref, False) < 1e-12
