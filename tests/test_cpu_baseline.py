"""The C++/OpenMP CPU baseline (oracle/heff_cpu.cpp) against the numpy oracle:
the four-step matvec to 1e-12 relative with the same block structure, and its
dgemm FLOP count equal to the oracle's count of the reference's schedule."""
import numpy as np
import pytest

from oracle import cpu_baseline as cb
from oracle.lanczos import eff_ham_mul_two_site_state, two_site_matvec_flops
from tests.workloads import synthetic_eff_ham


@pytest.mark.parametrize("kind,D,seed", [("heisenberg", 48, 11), ("hubbard", 40, 12), ("hubbard", 96, 13)])
def test_cpu_baseline_matches_oracle(kind, D, seed):
    eff, psi = synthetic_eff_ham(kind, D, seed=seed, nsect_half=3)
    plan = cb.CpuTwoSite(eff, psi)
    plan.matvec()
    got = plan.result()
    ref = eff_ham_mul_two_site_state(eff, psi)
    assert sorted(got.blocks) == sorted(ref.blocks)
    scale = max(float(np.abs(v).max()) for v in ref.blocks.values())
    err = max(float(np.abs(got.blocks[k] - ref.blocks[k]).max()) for k in ref.blocks)
    assert err <= 1e-12 * scale
    assert plan.flops == pytest.approx(two_site_matvec_flops(eff, psi), rel=1e-12)
    # a second state through the same plan
    x = np.random.default_rng(seed).normal(size=plan.psi_data.size)
    off = 0
    for k in sorted(psi.blocks):
        n = psi.blocks[k].size
        psi.blocks[k] = x[off:off + n].reshape(psi.blocks[k].shape)
        off += n
    plan.matvec(x)
    ref2 = eff_ham_mul_two_site_state(eff, psi)
    got2 = plan.result()
    err2 = max(float(np.abs(got2.blocks[k] - ref2.blocks[k]).max()) for k in ref2.blocks)
    assert err2 <= 1e-12 * max(float(np.abs(v).max()) for v in ref2.blocks.values())


def test_cpu_baseline_host_info():
    info = cb.host_info()
    assert info["nproc"] >= 1 and info["omp_threads"] >= 1 and info["blas"]
