"""Benchmark of the two-site effective-Hamiltonian matvec on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64]

Workload: U(1)xU(1) Hubbard 4x16 snake MPO (bond dimension w=18 in the
middle, d=4), bond dimension D=2000 with a Gaussian (N, 2Sz) sector profile,
seeded synthetic environments and state (seed 1003).  One step = one H_eff
matvec (three grouped-GEMM launches) with operands resident in HBM.  Value =
algorithmic matvec TFLOP/s: the sum of 2*M*N*K over the sector-GEMM pairs of
the reference's four-step schedule divided by the measured step time.

Prints one JSON line (the bench contract).  ``roofline`` describes the
dominant kernel (the stage with the longest average launch), timed with HIP
events on the stream it runs on: a contraction stage against the FP64 MFMA
peak with its share of the reference schedule's FLOPs, the MPO stage against
HBM bandwidth with its algorithmic bytes.  ``traffic`` is the HBM bytes per
launch of that kernel measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counters
(scaled by a copy kernel of known size) and recorded in
profiles/heff_traffic_hubbard_d2000.txt; it is null for other workloads.
``cpu_baseline`` times the C++/OpenMP CPU restatement of the same matvec
(block-sparse contractions with a host BLAS dgemm, oracle/heff_cpu.cpp) on this
host's cores over a bounded sample of the same workload.

    python bench.py --copy-calibration   # 1 GiB device copy, for counter scaling

Beside the FP64 matvec (``value``), the line carries the same matvec in the
split-BF16 6-term mode (``split``) and the 3-term mode (``split_alt``) and,
unless ``--no-sweep``, the sweep
comparison of the headline metric (``sweep``): the 4x16 Hubbard cylinder
(t=1, U=8, 56 electrons, Sz=0, the reference benchmark's initial product
state, ``benchmark/ultradmrg_hubbard2d.cpp:130-248``) is ramped in FP64 through
D = 10, 100, 1000 and then D=2000, up to 12 sweeps at each D (fewer once two
consecutive sweep energies agree to 1e-8), then two sweeps at D=2000 (Lanczos
1e-9 / 100, truncation 1e-8) are timed in FP64 and, from the same state, in the
split-BF16 6-term and 3-term modes; energies are compared relative to FP64.  Unless ``--no-chains``, ``batching`` compares 32 independent
random-field Heisenberg chains solved as one batch (one grouped launch per
matvec stage, hipGraph-captured Lanczos iterations) with a loop over the
chains; unless ``--no-tdvp``, ``tdvp`` times three complex128 TDVP steps at
D=1500 on the 4x8 Hubbard doublon quench; unless ``--no-chain``,
``heisenberg`` ramps the L=100 Heisenberg chain to D=1024 and times the
mid-chain matvec.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# The block SVD runs its eigen-decompositions on four streams beside the
# solver's own; with HIP's default of four hardware queues per process some
# of them share a queue and serialise (the largest bucket's lane then waits
# behind another's launches).  Eight queues keep them apart (0.25-0.4 s less
# SVD time per D=2000 sweep in a back-to-back comparison on one box).  Read
# by HIP at its first call; a value already in the environment wins.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eff-H matvec TFLOP/s + sweep sec at D=2000: split-BF16 vs FP64, |dE|<1e-8"
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
STAGES = ["stage1 L.psi", "stage2 MPO", "stage3 .R^T"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "heff_traffic_hubbard_d2000.txt")


def copy_calibration(nbytes=1 << 30, reps=5):
    """Plain device-to-device copy of a known size (scale for the HBM byte counters)."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "GBps": 2 * nbytes / el / 1e9}), flush=True)


def vector_kernels(D=2000, reps=50):
    """Krylov vector kernels on vectors of the D=2000 sizes, for a kernel
    trace (rocprofv3 --kernel-trace --stats): the library's dot, nrm2, axpy,
    scal and linear combination (k = 8) on the converged Hubbard state's
    two-site vector size (3,288,917 elements), ``reps`` calls each, then a
    30-iteration Lanczos solve on the synthetic D=2000 workload (its fused
    residual, check and normalisation kernels).  Prints the host-clock time
    per call (each call returns after completion) and the element counts the
    trace's durations are divided by."""
    import torch
    from ultradmrg_amd import tensor as T
    from ultradmrg_amd.solver import EffectiveHamiltonian, LanczosParams
    from ultradmrg_amd.workloads import synthetic_eff_ham
    n = 3288917
    g = torch.Generator(device="cuda").manual_seed(7)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    bases = [torch.randn(n, dtype=torch.float64, device="cuda", generator=g) for _ in range(8)]
    ops = {"dot": (lambda: T.dot(x, y), 2), "nrm2": (lambda: T.nrm2(x), 1),
           "axpy": (lambda: T.axpy(1e-3, x, y), 3), "scal": (lambda: T.scal(1.0, x), 2),
           "linear_combine_k8": (lambda: T.linear_combine([0.1] * 8, bases, 0.0, y), 9)}
    out = {"n": n, "element_bytes": 8, "per_call": {}}
    for name, (fn, touches) in ops.items():
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        us = (time.perf_counter() - t0) / reps * 1e6
        out["per_call"][name] = {"us": us, "bytes": touches * n * 8, "GBps_per_call": touches * n * 8 / us / 1e3}
    d_eff, x0 = synthetic_eff_ham("hubbard", D, seed=1003)
    H = EffectiveHamiltonian(d_eff, x0)
    res = H.lanczos(x0, LanczosParams(1e-30, 30))
    torch.cuda.synchronize()
    out["lanczos"] = {"n": int(x0.data.numel()), "iterations": res.iters}
    print(json.dumps(out), flush=True)


def svd_profile(D=2000):
    """Phase times of the block SVD over one FP64 D=2000 sweep of the 4x16
    Hubbard cylinder (after a short ramp: two sweeps at D = 10, 100, 1000,
    three at D=2000).  The library's QLG_SVD_PROFILE=1 mode drains the stream
    at every phase end and each eigensolver lane at its end, and prints one
    line per call; they are read back from the process's stderr and summed:
    per phase, the eigensolver lanes' finishing times, and the sector shapes."""
    import re
    import tempfile
    import torch
    from ultradmrg_amd import mpo as M
    from ultradmrg_amd.solver import LanczosParams
    from ultradmrg_amd.sweep import SweepParams, TwoSiteVMPS, direct_state_mps
    m = M.hubbard_cylinder(16, 4, t=1.0, U=8.0)
    per_hole, labs, sz = m.n // 8, [], 0
    for i in range(m.n):
        if i % per_hole == per_hole - 1:
            labs.append(3)
        else:
            labs.append(sz % 2 + 1)
            sz += 1
    lp = LanczosParams(1e-9, 100)
    drv = TwoSiteVMPS(direct_state_mps(m, labs), M.build_mpo(m), SweepParams(1, 10, 10, 1e-8, lp, verbose=False))
    for Dr, ns in ((10, 2), (100, 2), (1000, 2), (D, 3)):
        drv.p = SweepParams(1, Dr, Dr, 1e-8, lp, verbose=False)
        for k in range(ns):
            e = drv.sweep()
            print(f"[bench] svd profile ramp D={Dr} sweep {k + 1}: E = {e:.10f}", file=sys.stderr, flush=True)
    st = drv.stats
    for k in ("lanczos_s", "svd_s", "env_s", "contract_s", "plan_s"):
        setattr(st, k, 0)
    torch.cuda.synchronize()
    sys.stderr.flush()
    with tempfile.TemporaryFile(mode="w+") as tmp:
        saved = os.dup(2)
        os.dup2(tmp.fileno(), 2)
        os.environ["QLG_SVD_PROFILE"] = "1"
        os.environ["QLG_VERBOSE"] = "2"   # the planner's host phase times
        try:
            e = drv.sweep()
            torch.cuda.synchronize()
        finally:
            os.environ.pop("QLG_SVD_PROFILE", None)
            os.environ.pop("QLG_VERBOSE", None)
            sys.stderr.flush()
            os.dup2(saved, 2)
            os.close(saved)
        tmp.seek(0)
        lines = tmp.read().splitlines()
    plans = [l for l in lines if l.startswith("[qlg plan ms]")]
    plan_ms = {}
    for l in plans:
        for k, v in re.findall(r" ([a-z0-9+-]+) ([\d.]+)", l.split("]", 1)[1]):
            plan_ms[k] = plan_ms.get(k, 0.0) + float(v)
    calls = [l for l in lines if l.startswith("[qlg svd prof]")]
    lanes = [l for l in lines if l.startswith("[qlg eigh prof]")]
    phases = {}
    nmax = []
    for l in calls:
        for k, v in re.findall(r" (gram|eigh|project|refine|orthonormalise|assemble) ([\d.]+)", l):
            phases[k] = phases.get(k, 0.0) + float(v) / 1e3
        mm = re.search(r"n_max (\d+)", l)
        if mm:
            nmax.append(int(mm.group(1)))
    lane_max = [max(float(x) for x in re.findall(r"\] ([\d.]+) ms", l)) for l in lanes]
    lane_big = []
    for l in lanes:   # finishing time of the lane holding the largest matrices
        groups = re.findall(r"lane \d+ \[([^\]]*)\] ([\d.]+) ms", l)
        best = max(groups, key=lambda g: max(int(x.split("x")[0]) for x in g[0].split()))
        lane_big.append(float(best[1]))
    sweep_stats = {k: getattr(st, k) for k in ("lanczos_s", "svd_s", "env_s", "contract_s", "plan_s")}
    out = {"workload": f"Hubbard 16x4 cylinder, one FP64 sweep at D={D}", "energy": e, "svd_calls": len(calls),
           "sweep_s": st.sweep_seconds[-1], "sweep_phase_s": sweep_stats, "plans": len(plans),
           "plan_phase_ms_total": plan_ms,
           "svd_s": st.svd_s, "phase_s": phases, "largest_sector_order_median": sorted(nmax)[len(nmax) // 2],
           "eigh_lanes_critical_ms_median": sorted(lane_max)[len(lane_max) // 2] if lane_max else None,
           "eigh_largest_bucket_lane_ms_median": sorted(lane_big)[len(lane_big) // 2] if lane_big else None,
           "largest_call": max(calls, key=lambda l: int(re.search(r"n_max (\d+)", l).group(1))) if calls else None,
           "slowest_lanes": max(lanes, key=lambda l: max(float(x) for x in re.findall(r"\] ([\d.]+) ms", l)))
           if lanes else None}
    print(json.dumps(out), flush=True)


def recorded_traffic(D, precision, stage):
    """HBM bytes per launch of `stage` from the recorded counter runs (or None)."""
    if D != 2000 or precision != "fp64" or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        rec = json.load(f)
    v = rec.get("bytes_per_launch", {}).get(stage)
    return float(v) if v is not None else None


def captured_matvec(drv, site, modes, reps=20):
    """The reference profiler's measurement (profiler/single_processor_lanczos_performance_test.cc:76-119)
    on tensors captured from a running sweep: the two-site effective Hamiltonian
    at sites (site, site + 1) of ``drv``'s current state, timed in each
    precision mode, with per-stage kernel times and roofline fractions."""
    import torch
    from ultradmrg_amd.solver import EffectiveHamiltonian
    from ultradmrg_amd.tensor import contract
    n = len(drv.mps)
    for i in range(site):              # centre to (site, site + 1): left-canonicalise 0 .. site - 1
        drv.update("r", i)
    psi = contract(drv.mps[site], drv.mps[site + 1], 1)
    H = EffectiveHamiltonian([drv.lenvs[site], drv.mpo[site], drv.mpo[site + 1], drv.renvs[n - site - 2]], psi)
    x = H.adopt(psi)
    y = H.new_vector(zero=False)
    algo, _ = H.flops()
    s_algo, s_bytes = H.stage_work()
    bond = drv.mps[site].axes[2]
    out = {"site": [site, site + 1], "bond_dim": sum(bond.degs), "bond_sectors": bond.nsect,
           "largest_sector": max(bond.degs), "psi_elems": H.size, "algo_gflop_per_matvec": algo / 1e9}
    for mode in modes:
        H.set_precision(mode)
        ms = time_matvec(H, x, y, reps, 3)
        H.enable_timing(True)
        acc = [0.0, 0.0, 0.0]
        for _ in range(10):
            H.matvec_into(x.data, y.data, sync=True)
            acc = [p + q for p, q in zip(acc, H.stage_ms())]
        H.enable_timing(False)
        st = [v / 10 for v in acc]
        rec = {"ms_per_matvec": ms, "tflops": algo / (ms * 1e-3) / 1e12, "stage_ms": st,
               "stage_tflops": [f / (t * 1e-3) / 1e12 if t > 0 else None for f, t in zip(s_algo, st)],
               "mpo_stage_gbs": s_bytes[1] / (st[1] * 1e-3) / 1e9 if st[1] > 0 else None}
        if mode == "fp64":
            rec["frac_of_fp64_peak"] = rec["tflops"] / FP64_PEAK_TFLOPS
        else:
            mult = {"bf16x6": 6, "bf16x3": 3}.get(mode, 1)
            rec["issued_frac_of_bf16_peak_stage3"] = s_algo[2] * mult / (st[2] * 1e-3) / 1e12 / BF16_PEAK_TFLOPS
        out[mode] = rec
    for mode in modes:
        if mode != "fp64":
            out[mode]["speedup_vs_fp64"] = out["fp64"]["ms_per_matvec"] / out[mode]["ms_per_matvec"]
    H.close()
    return out


def hubbard_sweep_comparison(split="bf16x6", alt=(), lx=16, ramp=(10, 100, 1000), D=2000, max_converge=12,
                             conv_tol=1e-8, nsweeps=2, with_matvec=True, dmrg_form=True):
    """The D=2000 sweep workload: the 4x16 Hubbard cylinder (t=1, U=8, 56 electrons, Sz=0, the
    reference benchmark's initial product state, benchmark/ultradmrg_hubbard2d.cpp:130-248;
    Lanczos 1e-9 / 100, truncation 1e-8) ramped through the benchmark's bond
    dimensions ``ramp`` = 10, 100, 1000 and then D=2000 (:237-248), at each D
    sweeping until two consecutive sweep energies agree to ``conv_tol``
    relative or ``max_converge`` sweeps are done (the benchmark runs 12 per D).  From that converged state ``nsweeps`` sweeps are timed in FP64, in the
    split mode and in the ``alt`` modes; then one sweep per precision through
    the DMRG-form executor (FiniteDMRG's path, MatReprMPO and block-operator
    groups, ``dmrg_form``); the matvec at the chain centre is then timed on the
    captured converged tensors (``matvec_real``)."""
    import torch
    from ultradmrg_amd import mpo as M
    from ultradmrg_amd.solver import LanczosParams
    from ultradmrg_amd.sweep import SweepParams, TwoSiteVMPS, direct_state_mps

    m = M.hubbard_cylinder(lx, 4, t=1.0, U=8.0)
    n = m.n
    per_hole, labs, sz = n // 8, [], 0
    for i in range(n):   # one hole per 8 sites, alternating spins elsewhere
        if i % per_hole == per_hole - 1:
            labs.append(3)
        else:
            labs.append(sz % 2 + 1)
            sz += 1
    lp = LanczosParams(1e-9, 100)
    drv = TwoSiteVMPS(direct_state_mps(m, labs), M.build_mpo(m),
                      SweepParams(1, ramp[0], ramp[0], 1e-8, lp, verbose=False))
    t0 = time.perf_counter()

    def progress(what, e):
        print(f"[bench] {what}: E = {e:.10f} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    ramp_sweeps = []
    for Dr in ramp:
        drv.p = SweepParams(1, Dr, Dr, 1e-8, lp, verbose=False)
        es = [drv.sweep()]
        progress(f"ramp D={Dr} sweep 1", es[-1])
        while len(es) < max_converge:
            es.append(drv.sweep())
            progress(f"ramp D={Dr} sweep {len(es)}", es[-1])
            if abs(es[-1] - es[-2]) <= conv_tol * abs(es[-1]):
                break
        ramp_sweeps.append(len(es))
    drv.p = SweepParams(1, D, D, 1e-8, lp, verbose=False)
    conv = [drv.sweep()]
    progress(f"D={D} sweep 1", conv[-1])
    while len(conv) < max_converge:
        conv.append(drv.sweep())
        progress(f"D={D} sweep {len(conv)}", conv[-1])
        if abs(conv[-1] - conv[-2]) <= conv_tol * abs(conv[-1]):
            break
    ramp_s = time.perf_counter() - t0
    saved = (list(drv.mps), list(drv.lenvs), list(drv.renvs))   # updates replace tensors, never mutate them
    out = {"model": f"Hubbard {lx}x4 cylinder, U=8, t=1, {n - n // per_hole} electrons, Sz=0",
           "ramp": list(ramp), "ramp_sweeps": ramp_sweeps, "D": D, "converge_sweeps": len(conv),
           "converge_energies": conv,
           "converged_rel_change": abs(conv[-1] - conv[-2]) / abs(conv[-1]) if len(conv) > 1 else None,
           "ramp_and_converge_s": ramp_s, "timed_sweeps": nsweeps}
    for prec in ("fp64", split) + tuple(alt):
        drv.mps, drv.lenvs, drv.renvs = list(saved[0]), list(saved[1]), list(saved[2])
        drv.p = SweepParams(1, D, D, 1e-8, lp, precision=prec, verbose=False)
        st = drv.stats
        for k in ("lanczos_s", "svd_s", "env_s", "contract_s", "plan_s", "matvecs", "algo_flops"):
            setattr(st, k, 0)
        secs, es = [], []
        for _ in range(nsweeps):
            es.append(drv.sweep())
            secs.append(drv.stats.sweep_seconds[-1])
            progress(f"timed {prec} sweep", es[-1])
        torch.cuda.synchronize()
        out[prec] = {"sweep_s": secs, "energy": es[-1], "sweep_energies": es, "lanczos_s": st.lanczos_s, "svd_s": st.svd_s,
                     "env_s": st.env_s, "contract_s": st.contract_s, "plan_s": st.plan_s, "matvecs": st.matvecs,
                     "lanczos_tflops": st.algo_flops / max(st.lanczos_s, 1e-9) / 1e12}
    e64, es = out["fp64"]["energy"], out[split]["energy"]
    out["split"] = split
    out["speedup"] = sum(out["fp64"]["sweep_s"]) / sum(out[split]["sweep_s"])
    out["rel_dE"] = abs(es - e64) / abs(e64)
    out["alt"] = {m: {"speedup": sum(out["fp64"]["sweep_s"]) / sum(out[m]["sweep_s"]),
                      "rel_dE": abs(out[m]["energy"] - e64) / abs(e64)} for m in alt}
    if dmrg_form:
        from ultradmrg_amd.dmrg import DMRGExecutor, mat_repr_mpo
        mro = mat_repr_mpo(drv.mpo)
        out["dmrg_form"] = {}
        for prec in ("fp64", split):
            t = time.perf_counter()
            ex = DMRGExecutor(list(saved[0]), mro, SweepParams(1, D, D, 1e-8, lp, precision=prec, verbose=False))
            init_s = time.perf_counter() - t
            e = ex.sweep()
            progress(f"DMRG-form {prec} sweep", e)
            st = ex.stats
            out["dmrg_form"][prec] = {"sweep_s": st.sweep_seconds, "block_op_init_s": init_s, "energy": e,
                                      "lanczos_s": st.lanczos_s, "svd_s": st.svd_s, "block_op_update_s": st.env_s,
                                      "contract_s": st.contract_s, "plan_s": st.plan_s, "matvecs": st.matvecs}
        d64 = out["dmrg_form"]["fp64"]
        d64["rel_dE_vs_vmps_form"] = abs(d64["energy"] - out["fp64"]["sweep_energies"][0]) / abs(d64["energy"])
        out["dmrg_form"]["speedup"] = d64["sweep_s"][0] / out["dmrg_form"][split]["sweep_s"][0]
        out["dmrg_form"]["rel_dE"] = abs(out["dmrg_form"][split]["energy"] - d64["energy"]) / abs(d64["energy"])
    # the matvec on the converged state's centre tensors
    if with_matvec:
        drv.mps, drv.lenvs, drv.renvs = list(saved[0]), list(saved[1]), list(saved[2])
        drv.p = SweepParams(1, D, D, 1e-8, lp, verbose=False)
        out["matvec_real"] = captured_matvec(drv, n // 2 - 1, ("fp64", split) + tuple(alt))
    return out


def heisenberg_ramp_measurement(L=100, ramp=(256, 384, 512, 768, 1024, 1024), seed=1002, reps=30):
    """The plain FP64 form on the Heisenberg chain: L=100, J=1, Sz=0, a seeded
    random MPS (seed 1002) swept through the D ramp 256 -> 1024 (one sweep
    each, Lanczos 1e-10 / 100, truncation 1e-10).  Then the centre is moved to
    site 49 and the two-site matvec at sites (49, 50) is timed, as the
    reference's profiler captures a mid-chain site
    (``profiler/single_processor_lanczos_performance_test.cc:76-119``)."""
    import torch
    from ultradmrg_amd import mpo as M
    from ultradmrg_amd.solver import EffectiveHamiltonian, LanczosParams
    from ultradmrg_amd.sweep import SweepParams, TwoSiteVMPS, random_mps
    from ultradmrg_amd.tensor import contract

    m = M.heisenberg_chain(L)
    lp = LanczosParams(1e-10, 100)
    t0 = time.perf_counter()
    drv = TwoSiteVMPS(random_mps(m, (0,), ramp[0], seed=seed), M.build_mpo(m),
                      SweepParams(1, ramp[0], ramp[0], 1e-10, lp, verbose=False))
    energies, secs = [], []
    for Dr in ramp:
        drv.p = SweepParams(1, Dr, Dr, 1e-10, lp, verbose=False)
        energies.append(drv.sweep())
        secs.append(drv.stats.sweep_seconds[-1])
        print(f"[bench] Heisenberg L={L} sweep D={Dr}: E = {energies[-1]:.10f} ({secs[-1]:.1f} s, "
              f"SVD {drv.stats.svd_s:.1f} s so far)", file=sys.stderr, flush=True)
    total_s = time.perf_counter() - t0
    mid = L // 2 - 1
    for i in range(mid):
        drv.update("r", i)
    psi = contract(drv.mps[mid], drv.mps[mid + 1], 1)
    H = EffectiveHamiltonian([drv.lenvs[mid], drv.mpo[mid], drv.mpo[mid + 1], drv.renvs[L - mid - 2]], psi)
    x = H.adopt(psi)
    y = H.new_vector(zero=False)
    ms = time_matvec(H, x, y, reps, 3)
    algo, _ = H.flops()
    s_algo, _ = H.stage_work()
    H.enable_timing(True)
    acc = [0.0, 0.0, 0.0]
    for _ in range(10):
        H.matvec_into(x.data, y.data, sync=True)
        acc = [p + q for p, q in zip(acc, H.stage_ms())]
    H.enable_timing(False)
    st_ms = [v / 10 for v in acc]
    bond = drv.mps[mid].axes[2]
    out = {"model": f"Heisenberg chain L={L}, J=1, Sz=0", "ramp": list(ramp), "sweep_s": secs,
           "sweep_energies": energies, "total_s": total_s, "site": [mid, mid + 1], "bond_dim": sum(bond.degs),
           "bond_sectors": bond.nsect, "largest_sector": max(bond.degs), "psi_elems": H.size,
           "algo_gflop_per_matvec": algo / 1e9, "matvec_ms": ms, "matvec_tflops": algo / (ms * 1e-3) / 1e12,
           "stage_ms": st_ms,
           "stage_tflops": [f / (t * 1e-3) / 1e12 if t > 0 else None for f, t in zip(s_algo, st_ms)],
           "frac_of_fp64_peak": algo / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    H.close()
    return out


def tdvp_measurement(lx=8, ly=4, D=1500, ramp=(64, 256, 512, 1000, 1500), steps=3, tau=0.05, split="bf16x6"):
    """The TDVP form on a bounded sample: the 4x8 Hubbard cylinder (t=1, U=8,
    32 electrons, Sz=0) ramped in FP64 DMRG through ``ramp`` (one sweep each),
    c^dag_up c^dag_dn applied on the centre site (the reference's cupccdnc),
    then ``steps`` complex128 two-site TDVP steps (tau=0.05, Krylov dimension
    20, D=1500), once with FP64 and once with split-BF16 sector GEMMs from the
    same state.  Reports seconds per step with the expmv / SVD / environment
    split, the densities' sum (34 electrons) as a conservation check, and the
    largest local-density difference between the two forms."""
    import numpy as np
    import torch
    from ultradmrg_amd import mpo as M
    from ultradmrg_amd import tdvp as T
    from ultradmrg_amd.solver import LanczosParams
    from ultradmrg_amd.sweep import SweepParams, TwoSiteVMPS, direct_state_mps

    m = M.hubbard_cylinder(lx, ly, t=1.0, U=8.0)
    n = m.n
    labs = [1 if i % 2 == 0 else 2 for i in range(n)]
    lp = LanczosParams(1e-8, 50)
    mpo = M.build_mpo(m)
    drv = TwoSiteVMPS(direct_state_mps(m, labs), mpo, SweepParams(1, ramp[0], ramp[0], 1e-8, lp, verbose=False))
    t0 = time.perf_counter()
    for Dr in ramp:
        drv.p = SweepParams(1, Dr, Dr, 1e-8, lp, verbose=False)
        drv.sweep()
    ramp_s = time.perf_counter() - t0
    e0 = drv.stats.log[-1]["e0"]
    mps = [T.to_complex(t) for t in drv.mps]
    site = n // 2
    T.apply_local_op(mps, m.ops["cupccdnc"], site)
    mps[0] = T.BlockTensor(mps[0].axes, mps[0].coords, mps[0].data / torch.linalg.vector_norm(mps[0].data),
                           mps[0].offsets)
    out = {"model": f"Hubbard {lx}x{ly} cylinder, U=8, t=1, half filling, doublon on site {site}",
           "ramp": list(ramp), "ramp_s": ramp_s, "ground_energy": e0, "D": D, "tau": tau, "krylov_dim": 20,
           "steps": steps, "dtype": "c128"}
    dens = {}
    every = max(1, steps // 4)      # densities compared at four times over the run (every step for <= 4)
    for prec in ("fp64", split):
        td = T.TwoSiteTDVP(list(mps), mpo, T.TDVPParams(tau, steps, 1, D, 1e-8, LanczosParams(1e-10, 20),
                                                        precision=prec))
        dens[prec] = {}
        for k in range(1, steps + 1):
            td.sweep()
            if k % every == 0 or k == steps:
                dens[prec][k] = td.measure(m.ops["nf"])
            print(f"[bench] TDVP {prec} step {k}/{steps}: {td.stats.step_seconds[-1]:.2f} s", file=sys.stderr,
                  flush=True)
        st = td.stats
        last = dens[prec][steps]
        out[prec] = {"step_s": st.step_seconds, "expmv_s": st.expmv_s, "svd_s": st.svd_s, "env_s": st.env_s,
                     "krylov_matvecs": st.matvecs, "density_sum": float(np.sum(last)),
                     "centre_density": float(last[site])}
    out["split"] = split
    out["density_checked_at_steps"] = sorted(dens[split])
    out["max_density_error"] = max(float(np.abs(dens[split][k] - dens["fp64"][k]).max()) for k in dens[split])
    out["step_speedup"] = sum(out["fp64"]["step_s"]) / sum(out[split]["step_s"])
    return out


def chains_comparison(n=32, seed=1004, reps=20):
    """The batching form: n random-field Heisenberg chains (L ~ U[32, 128],
    fields U(-2, 2), D in {128, 256, 512, 768}, seed 1004), the centre two-site
    effective Hamiltonian of each.  Aggregate matvec TFLOP/s of the batch plan
    (one grouped launch per stage for all chains) against a loop over the
    chains' own plans, and the batched Lanczos solve (hipGraph-captured
    iterations, replayed) against separate solves, same energies."""
    import torch
    from ultradmrg_amd.chains import random_field_chains
    from ultradmrg_amd.solver import BatchedEffectiveHamiltonian, EffectiveHamiltonian, LanczosParams

    def sync_time():
        torch.cuda.synchronize()
        return time.perf_counter()

    t0 = time.perf_counter()
    chains = random_field_chains(n, seed)
    hams = [EffectiveHamiltonian(c.eff_ham, c.psi) for c in chains]
    build_s = time.perf_counter() - t0
    xs = [h.adopt(c.psi) for h, c in zip(hams, chains)]
    ys = [h.new_vector(zero=False) for h in hams]
    flops = sum(h.flops()[0] for h in hams)
    for h, x, y in zip(hams, xs, ys):          # settle tile shapes
        h.matvec_into(x.data, y.data, sync=False)
    t0 = sync_time()
    for _ in range(reps):
        for h, x, y in zip(hams, xs, ys):
            h.matvec_into(x.data, y.data, sync=False)
    loop_ms = (sync_time() - t0) * 1e3 / reps
    batch = BatchedEffectiveHamiltonian(hams, graphs=True)
    bx = batch.pack([c.psi for c in chains])
    by = torch.empty_like(bx)
    batch.matvec_into(bx, by)
    t0 = sync_time()
    for _ in range(reps):
        batch.matvec_into(bx, by, sync=False)
    batch_ms = (sync_time() - t0) * 1e3 / reps
    prm = LanczosParams(1e-10, 100)
    t0 = sync_time()
    single = [h.lanczos(c.psi, prm) for h, c in zip(hams, chains)]
    loop_lz = sync_time() - t0
    batch.lanczos([c.psi for c in chains], prm)          # captures the iteration graphs
    t0 = sync_time()
    res = batch.lanczos([c.psi for c in chains], prm)    # replays them
    batch_lz = sync_time() - t0
    rel = max(abs(r.gs_eng - s.gs_eng) / abs(s.gs_eng) for r, s in zip(res, single))
    same_iters = all(r.iters == s.iters for r, s in zip(res, single))
    out = {"chains": n, "lengths": [c.length for c in chains], "D": [c.D for c in chains],
           "build_s": build_s, "gflop_per_matvec": flops / 1e9,
           "batched_matvec_ms": batch_ms, "batched_tflops": flops / (batch_ms * 1e-3) / 1e12,
           "loop_matvec_ms": loop_ms, "loop_tflops": flops / (loop_ms * 1e-3) / 1e12,
           "matvec_speedup": loop_ms / batch_ms,
           "lanczos_batched_graph_s": batch_lz, "lanczos_loop_s": loop_lz, "lanczos_speedup": loop_lz / batch_lz,
           "lanczos_iters": [r.iters for r in res], "same_iterations": same_iters, "max_rel_dE": rel}
    batch.close()
    for h in hams:
        h.close()
    return out


def time_matvec(H, x, y, steps, warmup):
    H.ctx.synchronize()
    for _ in range(warmup):
        H.matvec_into(x.data, y.data, sync=False)
    H.ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        H.matvec_into(x.data, y.data, sync=False)
    H.ctx.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def cpu_baseline(eff, psi, min_runs=10, min_s=5.0, max_runs=200):
    """The C++/OpenMP restatement of the reference's four-step matvec with a
    host BLAS (oracle/heff_cpu.cpp) on the host cores: 2 warm-ups, then the
    median of at least ``min_runs`` matvecs (and at least ``min_s`` seconds)
    on the same D=2000 tensors."""
    import numpy as np
    from oracle import cpu_baseline as cb
    info = cb.host_info()
    plan = cb.CpuTwoSite(eff, psi)
    plan.matvec()
    plan.matvec()
    ts, t0 = [], time.perf_counter()
    while len(ts) < max_runs and (len(ts) < min_runs or time.perf_counter() - t0 < min_s):
        t = time.perf_counter()
        plan.matvec()
        ts.append(time.perf_counter() - t)
    med = float(np.median(ts))
    return {"value": plan.flops / med / 1e12, "unit": "TFLOP/s", "cores": info["omp_threads"], "kind": "port",
            "sample": f"median of {len(ts)} matvecs (after 2 warm-ups) of the same D=2000 tensors: C++/OpenMP "
                      f"block-sparse contractions, dgemm from {info['blas']}; host {info['cpu_model']}, "
                      f"nproc {info['nproc']}",
            "ms_per_matvec": med * 1e3, "nproc": info["nproc"], "cpu_model": info["cpu_model"], "blas": info["blas"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--D", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--svd-profile", action="store_true",
                    help="only the block SVD's phase times over one D=2000 sweep (QLG_SVD_PROFILE)")
    ap.add_argument("--vector-kernels", action="store_true",
                    help="only the Krylov vector kernels at the D=2000 sizes (for a kernel trace)")
    ap.add_argument("--split", default="bf16x6", help="reduced-precision mode compared with FP64")
    ap.add_argument("--split-alt", default="bf16x3",
                    help="second reduced-precision mode measured beside it in the matvec and sweep ('none': skip)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-chains", action="store_true")
    ap.add_argument("--no-tdvp", action="store_true")
    ap.add_argument("--no-chain", action="store_true", help="skip the L=100 Heisenberg ramp")
    ap.add_argument("--tdvp-steps", type=int, default=3, help="TDVP steps timed per form (the full D=1500 doublon-quench run is 40)")
    ap.add_argument("--sweep-profile", action="store_true",
                    help="only the Hubbard ramp and one D=2000 sweep per precision (for a profiler run)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this library does not shard the matvec across GPUs; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.vector_kernels:
        vector_kernels()
        return
    if a.svd_profile:
        svd_profile()
        return
    if a.sweep_profile:
        # one sweep of each kind at D=2000 after the ramp, for a profiler run
        out = hubbard_sweep_comparison(split=a.split, alt=(), max_converge=1, nsweeps=1, with_matvec=False,
                                       dmrg_form=False)
        print(json.dumps({"sweep_profile": out}), flush=True)
        return

    import torch
    from ultradmrg_amd.solver import EffectiveHamiltonian
    from ultradmrg_amd.workloads import synthetic_eff_ham

    d_eff, x = synthetic_eff_ham("hubbard", a.D, seed=1003)
    H = EffectiveHamiltonian(d_eff, x, precision=a.precision)
    psi_host = x
    x = H.adopt(x)
    y = H.new_vector(zero=False)
    algo, execf = H.flops()
    sflops = H.stage_flops()
    s_algo, s_bytes = H.stage_work()
    torch.cuda.synchronize()
    ms = time_matvec(H, x, y, a.steps, a.warmup)
    # per-stage kernel times (HIP events on the plan's stream), averaged
    H.enable_timing(True)
    acc = [0.0, 0.0, 0.0]
    nt = max(5, min(a.steps, 20))
    for _ in range(nt):
        H.matvec_into(x.data, y.data, sync=True)
        s = H.stage_ms()
        acc = [p + q for p, q in zip(acc, s)]
    H.enable_timing(False)
    st_ms = [v / nt for v in acc]
    dom = max(range(3), key=lambda i: st_ms[i])
    sec = st_ms[dom] * 1e-3
    traffic = recorded_traffic(a.D, a.precision, STAGES[dom])
    if dom == 1:   # MPO stage: HBM-bound, algorithmic bytes
        ach = s_bytes[dom] / sec / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
    else:          # contraction stage: MFMA-bound, its share of the reference schedule's FLOPs
        ach = s_algo[dom] / sec / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP64_PEAK_TFLOPS}
    roof.update({"traffic": traffic, "kernel": STAGES[dom], "launch_ms": st_ms[dom],
                 "algo_gflop": s_algo[dom] / 1e9, "algo_bytes": s_bytes[dom]})
    out = {
        "metric": METRIC,
        "value": algo / (ms * 1e-3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else a.precision,
        "data": "synthetic (seeded block-sparse environments and state)",
        "config": {"workload": f"two-site H_eff matvec, Hubbard 4x16 U1xU1 MPO (w=18, d=4), D={a.D}",
                   "D": a.D, "psi_elems": H.size, "algo_gflop_per_matvec": algo / 1e9,
                   "exec_gflop_per_matvec": execf / 1e9, "precision": a.precision,
                   "stage_ms": st_ms, "stage_exec_gflop": [f / 1e9 for f in sflops],
                   "stage_algo_gflop": [f / 1e9 for f in s_algo], "stage_algo_mb": [b / 1e6 for b in s_bytes]},
        "roofline": roof,
    }
    # the same matvec with the split-precision sector GEMMs
    def split_line(mode):
        H.set_precision(mode)
        ms_split = time_matvec(H, x, y, a.steps, a.warmup)
        H.enable_timing(True)
        acc = [0.0, 0.0, 0.0]
        for _ in range(nt):
            H.matvec_into(x.data, y.data, sync=True)
            acc = [p + q for p, q in zip(acc, H.stage_ms())]
        H.enable_timing(False)
        H.set_precision(a.precision)
        mult = {"fp32": 1, "bf16x6": 6, "bf16x3": 3}.get(mode, 1)
        return {"precision": mode, "value": algo / (ms_split * 1e-3) / 1e12, "unit": "TFLOP/s",
                "ms_per_step": ms_split, "speedup_vs_fp64": ms / ms_split,
                "stage_ms": [v / nt for v in acc],
                "issued_mfma_tflops": execf * mult / (ms_split * 1e-3) / 1e12,
                "note": "algorithmic FLOPs per second; issued = executed FLOPs x products per term"}

    alt = () if a.split_alt in ("none", a.split, a.precision) else (a.split_alt,)
    if a.split != a.precision:
        out["split"] = split_line(a.split)
    for mode in alt:
        out["split_alt"] = split_line(mode)
    H.close()
    if not a.no_chain:
        out["heisenberg"] = heisenberg_ramp_measurement()
    if not a.no_chains:
        out["batching"] = chains_comparison()
    if not a.no_tdvp:
        out["tdvp"] = tdvp_measurement(split=a.split, steps=a.tdvp_steps)
    if not a.no_sweep:
        out["sweep"] = hubbard_sweep_comparison(split=a.split, alt=alt)
        out["matvec_real"] = out["sweep"].pop("matvec_real")
    if not a.no_cpu:
        from tests.convert import from_device
        out["cpu_baseline"] = cpu_baseline([from_device(t) for t in d_eff], from_device(psi_host))
        if "sweep" in out:
            # a full CPU sweep at D=2000 does not fit the bench's time; its
            # matvecs alone, at the measured CPU rate, bound it from below
            sw = out["sweep"]["fp64"]
            per_sweep = sw["matvecs"] / len(sw["sweep_s"])
            flops_per_mv = sw["lanczos_tflops"] * sw["lanczos_s"] * 1e12 / sw["matvecs"]
            out["cpu_baseline"]["sweep_matvec_lower_bound_s"] = (per_sweep * flops_per_mv /
                                                                 (out["cpu_baseline"]["value"] * 1e12))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
