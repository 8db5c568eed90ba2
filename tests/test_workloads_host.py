"""The benchmark's synthetic workload, built on the CPU by the package's own MPO
builder (``ultradmrg_amd.workloads``), is the tests' oracle-built workload
block for block: the same sectors, the same block lists and the same seeded
values, so the matvec the bench times is the one the GPU parity tests check
against the oracle."""
import numpy as np
import pytest

from tests.convert import from_device
from tests.workloads import synthetic_eff_ham as oracle_workload
from ultradmrg_amd.workloads import synthetic_eff_ham


@pytest.mark.parametrize("kind,D,seed,nh", [("heisenberg", 96, 5, 4), ("hubbard", 64, 7, 2),
                                            ("hubbard", 400, 1003, None)])
def test_package_workload_equals_oracle_workload(kind, D, seed, nh):
    eff, psi = synthetic_eff_ham(kind, D, seed, nsect_half=nh, device="cpu")
    oeff, opsi = oracle_workload(kind, D, seed=seed, nsect_half=nh)
    for t, o in zip(eff + [psi], oeff + [opsi]):
        got = from_device(t)
        assert got.indexes == o.indexes
        assert sorted(got.blocks) == sorted(o.blocks)
        for k in o.blocks:
            np.testing.assert_array_equal(got.blocks[k], o.blocks[k])
