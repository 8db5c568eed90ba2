"""GPU parity at the benchmark's own shapes.

The workload the bench times (Hubbard 4x16 U(1)xU(1) MPO, D=2000, seed 1003,
built by the package's ``ultradmrg_amd.workloads``) and a profile with the
larger sectors of a converged D=2000 state (25 sectors up to 332 states,
~31 GFLOP per matvec, seed 1005) are run through the HIP matvec and compared
with the oracle's four-step contraction on the same tensors.

Bars: block lists exact; FP64 values within 1e-12 relative (max-abs error over
max-abs value); split-BF16 within the stated error bars (relative Frobenius
error 2e-6 for the 6-term form, 2e-4 for the 3-term form); algorithmic FLOPs
equal to the oracle's count."""
import numpy as np
import pytest
import torch

from oracle.lanczos import eff_ham_mul_two_site_state, two_site_matvec_flops
from tests.convert import from_device

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd.solver import EffectiveHamiltonian
    from ultradmrg_amd.workloads import synthetic_eff_ham

RED_TOL = {"bf16x6": 2e-6, "bf16x3": 2e-4}
CASES = {"bench_d2000": dict(kind="hubbard", D=2000, seed=1003),
         "large_sectors_d2000": dict(kind="hubbard", D=2000, seed=1005, nsect_half=3, width=1.4)}
_CACHE = {}


def _case(name):
    if name not in _CACHE:
        eff, psi = synthetic_eff_ham(**CASES[name])
        oeff, opsi = [from_device(t) for t in eff], from_device(psi)
        ref = eff_ham_mul_two_site_state(oeff, opsi)
        _CACHE[name] = (eff, psi, oeff, opsi, ref)
    return _CACHE[name]


def _errors(out, ref):
    got = out.blocks()
    assert sorted(got) == sorted(ref.blocks), "block structure differs"
    amax = max(float(np.abs(ref.blocks[k] - got[k]).max()) for k in ref.blocks)
    den = max(float(np.abs(b).max()) for b in ref.blocks.values())
    num2 = sum(float(np.sum((ref.blocks[k] - got[k]) ** 2)) for k in ref.blocks)
    den2 = sum(float(np.sum(b ** 2)) for b in ref.blocks.values())
    return amax / den, (num2 / den2) ** 0.5


@pytest.mark.parametrize("name", list(CASES))
def test_d2000_matvec_fp64_vs_oracle(name):
    print(f"[case] {name} fp64 vs oracle", flush=True)
    eff, psi, oeff, opsi, ref = _case(name)
    H = EffectiveHamiltonian(eff, psi)
    assert H.layout_coords == sorted(ref.blocks)
    out = H @ psi
    rel, frob = _errors(out, ref)
    print(f"  max-abs rel {rel:.2e}  Frobenius rel {frob:.2e}", flush=True)
    assert rel < 1e-12
    algo, _ = H.flops()
    assert algo == pytest.approx(two_site_matvec_flops(oeff, opsi), rel=1e-12)
    H.close()


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("name", list(CASES))
def test_d2000_matvec_split_vs_oracle(name, prec):
    print(f"[case] {name} {prec} vs oracle", flush=True)
    eff, psi, _, _, ref = _case(name)
    H = EffectiveHamiltonian(eff, psi, precision=prec)
    out = H @ psi
    _, frob = _errors(out, ref)
    print(f"  Frobenius rel {frob:.2e} (bar {RED_TOL[prec]:.0e})", flush=True)
    assert frob < RED_TOL[prec]
    H.close()
