#!/usr/bin/env python3
"""Benchmark of the state-vector gate-sweep path on one MI355X.

Workload (default): the QFT (QSEplayground/QFT.cpp:9-22: H + cphase(pi/2^k)
ladder + final swaps, 480 gates) applied to a seeded random normalised
30-qubit complex128 state resident in HBM (16 GiB). One step = one QFT over the
whole state. Metric: amplitude updates per second (sum over gates of the
amplitudes each gate can change: N for H, N/4 for cphase, N/2 for swap).

The JSON line also carries:
  roofline      the dominant kernel family's algorithmic HBM bytes per launch
                divided by its mean launch time (HIP events recorded on the
                library's own stream during the timed steps), against 8 TB/s
  cpu_baseline  the reference simulator's AVX2+OpenMP backend
                (oracle/_ref/libqse_ref.so, built from the reference sources)
                timed on this host on the same QFT gate list, one thread per
                physical core, median of 3 runs after a warm-up; the restated
                scalar oracle on a 32-gate prefix when the reference build is
                absent
  qaoa          QAOA-24q gradients/s, adjoint vs parameter shift (when built)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from quantumsimevo_amd import capi  # noqa: E402
from quantumsimevo_amd import circuits as C  # noqa: E402

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(kernel: str, n: int, precision: str):
    """HBM bytes per launch of the timed kernel as measured by rocprofv3 --pmc
    FETCH_SIZE/WRITE_SIZE (calibrated on a copy kernel) for this workload,
    from the committed summary; None for workloads without one."""
    f = ROOT / "profiles" / f"{kernel}_qft{n}_{precision}_traffic.md"
    if not f.exists():
        return None, None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), str(f.relative_to(ROOT))
    return None, None


def host_cores() -> dict:
    """Physical cores available to this process: the (core, socket) pairs that
    `lscpu -p` lists for the CPUs of the process's affinity mask (sysfs topology
    when lscpu is missing), capped by a cgroup CPU quota when one is set."""
    import math
    import subprocess
    cpus = sorted(os.sched_getaffinity(0))
    pairs = set()
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line and not line.startswith("#"):
                c, core, sock = line.split(",")[:3]
                if int(c) in cpus:
                    pairs.add((sock, core))
        how = "lscpu -p"
    except (OSError, ValueError, subprocess.SubprocessError):
        for c in cpus:
            try:
                top = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
                pairs.add(((top / "physical_package_id").read_text().strip(), (top / "core_id").read_text().strip()))
            except OSError:
                pairs.add(("?", str(c)))
        how = "sysfs topology"
    phys = max(len(pairs), 1)
    quota = None
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    use = min(phys, max(1, math.floor(quota))) if quota else phys
    return {"physical": phys, "logical": len(cpus), "cgroup_cpus": quota, "threads": use, "source": how}


def cpu_baseline(n: int, ops: np.ndarray, reps: int = 3) -> dict:
    """The reference's AVX2+OpenMP backend (kind "reference", built from its
    sources into oracle/_ref) on the whole gate list: one thread per physical
    core (OMP_PROC_BIND=close, OMP_PLACES=cores), one untimed warm-up over the
    first 32 gates (page faults, thread start), then the median of `reps` timed
    replays, in a child process so the OpenMP settings apply from its start.
    Without the reference build: the restated scalar oracle (kind "port", one
    thread) on the first 32 gates."""
    import subprocess
    cores = host_cores()
    ref_path = ROOT / "oracle" / "_ref" / "libqse_ref.so"
    if ref_path.exists():
        env = dict(os.environ, OMP_NUM_THREADS=str(cores["threads"]), OMP_PROC_BIND="close", OMP_PLACES="cores")
        t0 = time.perf_counter()
        r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-child", str(n), "--cpu-reps", str(reps)],
                           env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"cpu baseline child failed: {r.stderr[-2000:]}")
        # the reference's StateVector prints its allocation time on stdout too
        times = next(json.loads(x) for x in r.stdout.splitlines() if x.startswith('{"times"'))["times"]
        secs = float(np.median(times))
        upd = C.amp_updates(ops, n)
        return {"value": upd / secs, "unit": "amp-updates/s", "cores": cores["threads"], "kind": "reference",
                "sample": f"the whole {len(ops)}-gate QFT-{n} gate list ({upd:.4e} amp-updates) on "
                          f"CircuitUnitaryOperationAVX (AVX2+OpenMP, built from the reference sources), "
                          f"{cores['threads']} threads = physical cores ({cores['source']}: {cores['physical']} "
                          f"physical / {cores['logical']} logical CPUs in the affinity mask, cgroup quota "
                          f"{cores['cgroup_cpus']}), OMP_PROC_BIND=close OMP_PLACES=cores, one warm-up over the "
                          f"first 32 gates, median of {reps} runs: {', '.join(f'{t:.2f}' for t in times)} s "
                          f"(child wall {time.perf_counter() - t0:.1f} s)",
                "seconds_median": secs, "seconds": times}
    sub = np.ascontiguousarray(ops[:32])
    upd = C.amp_updates(sub, n)
    lib = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    lib.orc_run_ops.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    t0 = time.perf_counter()
    lib.orc_run_ops(psi.ctypes.data, n, sub.ctypes.data, len(sub), None)
    secs = time.perf_counter() - t0
    return {"value": upd / secs, "unit": "amp-updates/s", "cores": 1, "kind": "port",
            "sample": f"first 32 gates of the {n}-qubit QFT gate list ({upd:.3e} amp-updates, {secs:.2f} s) on the "
                      f"restated scalar FP64 oracle (the reference build is absent)"}


def cpu_child(n: int, reps: int) -> int:
    """Child process of cpu_baseline: times the reference AVX2 backend."""
    ops = np.ascontiguousarray(C.qft(n).ops())
    lib = ctypes.CDLL(str(ROOT / "oracle" / "_ref" / "libqse_ref.so"))
    lib.ref_time_ops_repeat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.ref_time_ops_repeat.restype = ctypes.c_int
    times = np.zeros(reps)
    rc = lib.ref_time_ops_repeat(1, n, ops.ctypes.data, len(ops), 32, 1, reps, times.ctypes.data)
    if rc != 0:
        return 1
    print(json.dumps({"times": times.tolist()}), flush=True)
    return 0


def qaoa_leg(ctx, batch: int, shift_sets: int) -> dict:
    """QAOA Max-Cut, 24 qubits, seeded random 3-regular graph (36 edges), p = 8
    (16 parameters, uniform in [0, pi)), cost sum_edges (1 - Z_i Z_j)/2 in FP64:
    batched adjoint gradients vs plain per-parameter-shift on the same GPU."""
    n, p = 24, 8
    edges = C.random_regular_graph(n, 3, seed=2024)
    ops = C.qaoa_maxcut(n, edges, p).ops()
    masks, w = C.maxcut_observable(edges)
    rng = np.random.default_rng(7)
    params = rng.random((batch, 2 * p)) * np.pi
    # warm-up: two parameter sets show every pass structure twice, which starts their
    # compilation (coefficients read at run time); wait for it, as a training loop would
    t0 = time.perf_counter()
    ctx.adjoint_expect(n, capi.C128, ops, params[:2], masks, w)
    ctx.jit_wait()
    ctx.sync()
    t_warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    E, G = ctx.adjoint_expect(n, capi.C128, ops, params, masks, w)
    t_adj = time.perf_counter() - t0
    log(f"QAOA-24 adjoint: batch {batch} in {t_adj:.3f} s")
    occ = np.nonzero(ops["param"] >= 0)[0]
    t0 = time.perf_counter()
    Gs = np.zeros((shift_sets, 2 * p))
    for k in occ:
        for sgn in (+1.0, -1.0):
            sh = ops.copy()
            sh["theta"][k] += sgn * np.pi / 2
            Es = ctx.batch_expect(n, capi.C128, sh, params[:shift_sets], masks, w)
            Gs[:, int(ops[k]["param"])] += sgn * 0.5 * float(ops[k]["scale"]) * Es
    t_shift = time.perf_counter() - t0
    log(f"QAOA-24 parameter shift: {shift_sets} set(s), {2 * len(occ)} evaluations each, {t_shift:.2f} s")
    err = float(np.max(np.abs(G[:shift_sets] - Gs)))
    return {"workload": "QAOA Max-Cut 24q, 3-regular graph (36 edges), p=8, 16 params, complex128, FP64 cost",
            "batch": batch, "adjoint_grads_per_s": batch / t_adj, "shift_grads_per_s": shift_sets / t_shift,
            "speedup": (batch / t_adj) / (shift_sets / t_shift), "max_abs_grad_diff_vs_shift": err,
            "shift_evaluations_per_grad": int(2 * len(occ)), "mean_energy": float(np.mean(E)),
            "warmup_seconds_incl_compile": t_warm}


def noise_leg(ctx) -> dict:
    """Batched Pauli-noise trajectories: 20-qubit random layered circuit (one of
    H/RX/RY/RZ/phase per qubit per layer, CNOT brickwork), depth 30, complex64;
    512 trajectories whose X/Z frames come from the Noise draw process (seeds
    0..511, bit- and phase-flip probability 1e-3 per eligible gate); 10 000
    shots per trajectory sampled from the FP64 CDF of the one executed state
    (frames permute outcomes, Noise.cpp:34-44)."""
    from quantumsimevo_amd import _quantumsimevo as Q
    n, depth, ntraj, shots = 20, 30, 512, 10000
    ops = C.random_layered(n, depth, seed=2024).ops()
    rng = np.random.default_rng(11)
    u = rng.random((ntraj, shots))
    st = ctx.state(n, capi.C64)
    for _ in range(2):  # warm-up: the second run compiles the circuit's passes (run-time specialisation)
        st.set_basis(0)
        st.run(ops)
    st.sample_frames(np.zeros(1, dtype=np.uint64), u[:1, :100])
    ctx.sync()
    t0 = time.perf_counter()
    x, _z = Q.noise_frames(n, ops["kind"], ops["target"], np.arange(ntraj, dtype=np.uint32), 1e-3, 1.0)
    t1 = time.perf_counter()
    st.set_basis(0)
    st.run(ops)
    out = st.sample_frames(x, u)
    t2 = time.perf_counter()
    # complex64 error against FP64: the same circuit on a complex128 state (which the tests hold to 1e-12 of
    # the oracle), after the timed region
    psi = st.download()
    st.close()
    s128 = ctx.state(n, capi.C128)
    s128.run(ops)
    ref = s128.download()
    s128.close()
    fid = float(abs(np.vdot(ref, psi)) ** 2 / (np.vdot(psi, psi).real * np.vdot(ref, ref).real))
    inter = interleaved_noise(ctx, n, ops, ntraj, u)
    return {"workload": f"{ntraj} noise trajectories x {shots} shots, {n}-qubit random layered circuit depth {depth}, "
                        "complex64 state, FP64 CDF",
            "mode": "deferred: the drawn flips as one Pauli frame after the last gate (the reference's model, "
                    "Noise.cpp:34-44)", "seconds": t2 - t0, "frames_seconds": t1 - t0,
            "trajectories_per_s": ntraj / (t2 - t0), "shots_per_s": ntraj * shots / (t2 - t0),
            "distinct_frames": int(len(np.unique(x))), "outcomes_checksum": int(np.bitwise_xor.reduce(out.ravel())),
            "c64_max_amp_err_vs_fp64": float(np.max(np.abs(psi - ref))), "c64_infidelity_vs_fp64": 1.0 - fid,
            "sampling": "outcomes equal the sequential-CDF lower_bound on the complex64 amplitudes (bit-exact, "
                        "tests/test_gpu_configs.py)",
            "interleaved": inter}


def interleaved_noise(ctx, n, ops, ntraj, u) -> dict:
    """The same trajectories (seeds 0..ntraj-1, the same draws) in the
    interleaved mode: each drawn flip acts right after its gate, so every
    trajectory with a flip is a gate list of its own, run from |0...0> and
    sampled; trajectories with identical flips share one run. The tile shape
    is fixed (12 qubits, runs of 16 amplitudes) so that each distinct list is
    planned once without the shape search, and each runs on the generic pass
    kernel (a list that runs once is not worth compiling)."""
    keep = ctx.get_fusion(capi.C64)
    ctx.set_fusion(capi.C64, 12, 4)
    ctx.set_jit(capi.JIT_OFF)  # each distinct list runs once: no run-time compilation
    try:
        st = ctx.state(n, capi.C64)
        st.set_basis(0)
        st.run(ops)
        ctx.sync()
        t0 = time.perf_counter()
        trajs = C.interleaved_trajectories(n, ops, np.arange(ntraj, dtype=np.uint32), 1e-3, 1.0)
        t1 = time.perf_counter()
        groups = {}
        for j, t in enumerate(trajs):
            groups.setdefault(t.tobytes(), []).append(j)
        csum = 0
        for key, members in groups.items():
            st.set_basis(0)
            st.run(trajs[members[0]])
            for j in members:
                csum ^= int(np.bitwise_xor.reduce(st.sample(u[j])))
        ctx.sync()
        t2 = time.perf_counter()
        st.close()
    finally:
        ctx.set_fusion(capi.C64, *keep)
        ctx.set_jit(capi.JIT_AUTO)
    return {"mode": "interleaved: each drawn flip right after its gate (no counterpart in the reference)",
            "seconds": t2 - t0, "insertions_seconds": t1 - t0, "distinct_runs": len(groups),
            "trajectories_per_s": ntraj / (t2 - t0), "shots_per_s": ntraj * u.shape[1] / (t2 - t0),
            "outcomes_checksum": csum,
            "check": "each trajectory's state and samples against the oracle's replay of its gate list "
                     "(tests/test_gpu_noise_modes.py)"}


def uneven_batch_leg(ctx) -> dict:
    """1024 ansatz circuits (RY+RZ per qubit per layer, linear CNOT ladder),
    qubits uniform in 10..22, depth uniform in 4..12, angles uniform in
    [0, 2pi), observable sum_i Z_i: energies and adjoint gradients in two
    persistent launches (forward programs, energies and observables; then the
    reverse sweeps as joint psi/lambda passes with in-tile generator probes),
    complex64 with FP64 accumulation; the same batch in complex128 gives the
    FP64 values the errors are stated against."""
    circs = C.uneven_ansatz_batch(1024, seed=2024, n_range=(10, 22), depth_range=(4, 12))
    masks = [1 << q for q in range(22)]
    w = [1.0] * 22
    small = circs[:8]
    ctx.batch_run_small(capi.C64, small, masks, w)  # warm-up (device buffer, code objects)
    # tile shape and team size of the engine measured on the batch itself (the team size
    # depends on how many large circuits there are), kept for the timed run
    t0 = time.perf_counter()
    tune = ctx.batch_autotune(capi.C64, circs, masks, w)
    tune_s = time.perf_counter() - t0
    ctx.reset_timing()
    ctx.enable_timing(True)
    t0 = time.perf_counter()
    E64, G64 = ctx.batch_run_small(capi.C64, circs, masks, w)
    t1 = time.perf_counter()
    ctx.enable_timing(False)
    kern = ctx.timing(capi.FAM_BATCH)
    E, Gr = ctx.batch_run_small(capi.C128, circs, masks, w)
    t2 = time.perf_counter()
    nparams = sum(len(c[2]) for c in circs)
    # independent FP64 values for a sample (the 8 largest circuits and 8 seeded others): the host-driven
    # complex128 adjoint, a different code path from the engine (tests compare all 1024 and the oracle)
    order = sorted(range(len(circs)), key=lambda i: (-circs[i][0], -len(circs[i][1])))
    pick = order[:8] + [int(i) for i in np.random.default_rng(3).choice(order[8:], 8, replace=False)]
    ctx.set_jit(capi.JIT_OFF)
    eh = gh = e64h = g64h = 0.0
    for i in pick:
        nq, ops_i, par_i = circs[i]
        zm, zw = C.zsum_observable(nq)
        Eh, Gh = ctx.adjoint_expect(nq, capi.C128, ops_i, par_i, zm, zw)
        eh, gh = max(eh, abs(E[i] - Eh[0])), max(gh, float(np.max(np.abs(Gr[i] - Gh[0]))))
        e64h, g64h = max(e64h, abs(E64[i] - Eh[0])), max(g64h, float(np.max(np.abs(G64[i] - Gh[0]))))
    ctx.set_jit(capi.JIT_AUTO)
    return {"workload": "1024 ansatz circuits, 10..22 qubits, depth 4..12, sum of Z, energies + adjoint gradients, "
                        "complex64 / FP64 accumulation, two persistent launches (forward; reverse as joint "
                        "psi/lambda passes with in-tile generator probes)",
            "seconds": t1 - t0, "kernel_seconds": kern["ms"] / 1e3, "launches": kern["launches"],
            "algorithmic_bytes": kern["bytes"], "achieved_gbs": kern["bytes"] / kern["ms"] / 1e6,
            "gradient_vectors_per_s": len(circs) / (t1 - t0),
            "parameters": nparams, "complex128_seconds": t2 - t1,
            "autotuned_shape": {"tile_qubits": tune["tile_qubits"], "low_qubits": tune["low_qubits"],
                                "team_workgroups": tune["team_workgroups"],
                                "sample": "the whole batch, 16 timed runs", "tuning_seconds": tune_s},
            "c64_max_abs_energy_err_vs_c128_engine": float(np.max(np.abs(E64 - E))),
            "c64_max_abs_grad_err_vs_c128_engine": float(max(np.max(np.abs(a - b)) for a, b in zip(G64, Gr))),
            "host_adjoint_check": {"circuits": len(pick), "c128_energy_err": eh, "c128_grad_err": gh,
                                   "c64_energy_err": e64h, "c64_grad_err": g64h,
                                   "reference": "qse_adjoint_expect complex128 (host-driven fused sweeps), one "
                                                "call per circuit"}}


def exchange_leg(ctx, n: int = 30) -> dict:
    """Pair exchange for gates on the lowest qubits, measured both ways
    (qse_exchange_probe): reps layers of RY on qubits 0..5 in one pass over a
    30-qubit complex128 state, amplitude pairs brought together through LDS
    between register layouts (the fused pass's way) or across wavefront lanes
    by DPP / ds_swizzle / ds_bpermute. Each figure is the second of two
    launches (HIP events); the rates count one read and one write of the state."""
    st = ctx.state(n, capi.C128)
    st.set_random(3)
    rows = []
    nbytes = 2.0 * 16.0 * 2 ** n
    for reps in (1, 2, 4, 8, 16):
        row = {"reps": reps, "gates": 6 * reps}
        for method, name in ((0, "lds"), (1, "dpp")):
            st.exchange_probe(method, reps)
            ms = st.exchange_probe(method, reps)
            row[f"{name}_ms"] = ms
            row[f"{name}_gbs"] = nbytes / (ms * 1e-3) / 1e9
        row["faster"] = "lds" if row["lds_ms"] <= row["dpp_ms"] else "dpp"
        rows.append(row)
    st.close()
    return {"workload": f"reps x RY on qubits 0..5 in one pass, {n}-qubit complex128 state, tiles of 4096 contiguous "
                        "amplitudes, 256 threads x 16", "rows": rows,
            "profile": "profiles/k_exchange_lds_dpp_30q_c128_kernel_stats.md"}


def layered_leg(ctx, n: int = 30, layers: int = 20) -> dict:
    """The plain-method random circuit at 30 qubits complex128: per layer one
    of H/RX/RY/RZ/phase per qubit (angles uniform in [0, 2pi)), then a CNOT
    brickwork; executed as fused passes on a resident seeded random state.
    Per-pass HBM rate from the library's HIP events (2 x 16 x 2^n bytes per
    pass), mean of 3 runs after 2 warm-up runs (the second compiles)."""
    ops = C.random_layered(n, layers, seed=7).ops()
    st = ctx.state(n, capi.C128)
    st.set_random(11)
    info = ctx.plan_info(n, capi.C128, ops)
    for _ in range(2):
        st.run(ops)
    ctx.sync()
    ctx.reset_timing()
    ctx.enable_timing(True)
    t0 = time.perf_counter()
    for _ in range(3):
        st.run(ops)
    ctx.sync()
    wall = (time.perf_counter() - t0) / 3
    ctx.enable_timing(False)
    f = ctx.timing(capi.FAM_FUSED)
    norm_err = abs(st.norm2() - 1.0)
    st.close()
    pass_ms = f["ms"] / max(f["launches"], 1)
    gbs = 2.0 * 16 * 2.0 ** n / (pass_ms * 1e-3) / 1e9
    log(f"layered {n}q x {layers}: {wall * 1e3:.1f} ms per run, {f['launches'] // 3} passes, {pass_ms:.3f} ms each "
        f"= {gbs:.0f} GB/s")
    return {"workload": f"random layered circuit, {n} qubits complex128, {layers} layers ({len(ops)} gates)",
            "ms_per_run": wall * 1e3, "passes_per_run": f["launches"] // 3, "tile_qubits": info["tile_bits"],
            "contiguous_low_qubits": info["low_bits"], "pass_mean_ms": pass_ms, "pass_gbs": gbs,
            "pass_frac_of_peak": gbs / 8000.0, "norm_err_after_3_runs": norm_err}


def dense_block_leg(ctx, restore_jit: int, n: int = 30) -> dict:
    """Matrix-core dense blocks against the fused gate path: a random layered
    circuit of `layers` layers on qubits 0..m-1 (one of H/RX/RY/RZ/phase per
    qubit per layer, CNOT brickwork) applied to an n-qubit complex128 state,
    either gate by gate in one fused pass (run-time specialised kernel) or as
    its composed 2^m x 2^m unitary with FP64 MFMA (qse_apply_dense). Kernel
    times are HIP events of the library (mean of 3 launches after 2 warm-up
    runs); both paths move 2 x 16 x 2^n bytes. The results of the two paths
    are compared on a 20-qubit state."""
    rows = []
    st = ctx.state(n, capi.C128)
    st.set_random(7)
    a, b = ctx.state(20, capi.C128), ctx.state(20, capi.C128)
    ctx.set_jit(capi.JIT_ALWAYS)
    for m in (4, 5, 6):
        for layers in (1, 16, 64):
            ops = C.random_layered(m, layers, seed=layers).ops()
            u = capi.compose_dense(m, ops)
            ctx.set_dense_blocks(0)  # the fused gate path, for the comparison
            for _ in range(2):
                st.run(ops)
                st.apply_dense(m, u)
            ctx.sync()
            ctx.reset_timing()
            ctx.enable_timing(True)
            for _ in range(3):
                st.run(ops)
            for _ in range(3):
                st.apply_dense(m, u)
            ctx.sync()
            ctx.enable_timing(False)
            f, d = ctx.timing(capi.FAM_FUSED), ctx.timing(capi.FAM_DENSE)
            # qse_run with its default routing: which path it takes for this block
            ctx.set_dense_blocks(20)
            ctx.reset_timing()
            ctx.enable_timing(True)
            st.run(ops)
            ctx.sync()
            ctx.enable_timing(False)
            routed = "mfma" if ctx.timing(capi.FAM_DENSE)["launches"] else "fused"
            ctx.set_dense_blocks(0)
            a.set_random(3)
            b.set_random(3)
            a.run(ops)
            b.apply_dense(m, u)
            diff = float(np.max(np.abs(a.download() - b.download())))
            ctx.set_dense_blocks(20)
            fms, dms = f["ms"] / max(f["launches"], 1), d["ms"] / max(d["launches"], 1)
            rows.append({"m": m, "layers": layers, "gates": int(len(ops)), "fused_ms": fms, "dense_mfma_ms": dms,
                         "dense_gbs": 2.0 * 16 * 2.0 ** n / (dms * 1e-3) / 1e9, "faster": "mfma" if dms < fms else "fused",
                         "qse_run_takes": routed, "max_abs_diff_20q": diff})
            log(f"dense block m={m} layers={layers}: fused {fms:.3f} ms, MFMA {dms:.3f} ms")
    ctx.set_jit(restore_jit)
    for s_ in (st, a, b):
        s_.close()
    return {"workload": f"random layered block on qubits 0..m-1 of a {n}-qubit complex128 state: fused gate pass "
                        "vs composed 2^m x 2^m unitary on the FP64 matrix cores", "rows": rows}


def qft_leg(ctx, args, n: int, prec: int, ops: np.ndarray, upd_step: float):
    """The headline line: QFT on a resident seeded random state (see the module docstring)."""
    plan0 = ctx.plan_info(n, prec, ops) if not args.unfused else {"low_bits": 0}
    # copy probe of this device for the pass's access pattern (tiled out-of-place copy of the state's size,
    # plain loads and stores; complex64 runs of 2^L amplitudes are 2^(L-1) 16-byte elements): a reference
    # point, not a bound (the passes use non-temporal accesses and paired tiles and run faster)
    # (buffers of the state's size, at most 2^31 16-byte elements = 32 GiB each)
    log2_el = min(n if prec == capi.C128 else n - 1, 31)
    run_el = max(plan0["low_bits"] - (0 if prec == capi.C128 else 1), 0)
    copy_contig = ctx.copy_bandwidth(log2_el, 12)
    copy_runs = ctx.copy_bandwidth(log2_el, run_el) if run_el < 12 else copy_contig
    log(f"copy probe: contiguous {copy_contig:.0f} GB/s, runs of 2^{run_el} x 16 B {copy_runs:.0f} GB/s")
    st = ctx.state(n, prec)
    tune = None
    if args.autotune and not args.unfused:
        tune = ctx.autotune(st, ops, candidates=3)
        log(f"autotune: {tune}")
    st.set_random(12345)
    plan = ctx.plan_info(n, prec, ops) if not args.unfused else {"passes": len(ops), "tile_bits": 0, "low_bits": 0,
                                                                  "stages": 0}
    passes = plan["passes"]
    log(f"QFT-{n} {args.precision}: {len(ops)} gates, {upd_step:.4e} amp-updates/step, plan {plan}")

    fused = not args.unfused
    for i in range(args.warmup):
        st.run(ops, fused=fused)
        ctx.sync()
        log(f"warmup {i + 1}/{args.warmup} done")

    ctx.reset_timing()
    ctx.enable_timing(True)
    ctx.sync()
    jit0 = ctx.jit_stats()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st.run(ops, fused=fused)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.enable_timing(False)
    jit1 = ctx.jit_stats()
    jit_timed = jit1["launches"] > jit0["launches"]
    if jit1["compiled"] != jit0["compiled"]:
        raise SystemExit("a program was compiled inside the timed region: raise --warmup")
    ms_step = (t1 - t0) * 1e3 / args.steps
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step")

    fams = {f: ctx.timing(f) for f in (capi.FAM_SWEEP, capi.FAM_FUSED)}
    dom_f = max(fams, key=lambda f: fams[f]["ms"])
    dom = fams[dom_f]
    mean_ms = dom["ms"] / max(dom["launches"], 1)
    bytes_per_launch = dom["bytes"] / max(dom["launches"], 1)
    achieved = bytes_per_launch / (mean_ms * 1e-3) if mean_ms > 0 else 0.0
    kname = {capi.FAM_SWEEP: "k_pairs (one-gate sweep)",
             capi.FAM_FUSED: "qse_jit_pass* (LDS-tiled pass, run-time specialised)" if jit_timed
             else "k_fused (LDS-tiled pass)"}[dom_f]
    log(f"dominant kernel {kname}: {dom['launches']} launches, mean {mean_ms:.4f} ms, "
        f"{bytes_per_launch / 1e9:.3f} GB/launch -> {achieved / 1e9:.1f} GB/s")

    prof_kernel = None
    if fused and dom_f == capi.FAM_FUSED:
        prof_kernel = "qse_jit_pass" if jit_timed else "k_fused"
    traffic, traffic_src = profiled_traffic(prof_kernel, n, args.precision) if prof_kernel else (None, None)
    # size-independent sanity: the norm stays 1 after K unitary steps
    nrm = st.norm2()
    if abs(nrm - 1.0) > (1e-9 if prec == capi.C128 else 1e-3):
        raise SystemExit(f"norm drifted: {nrm}")

    result = {
        "metric": f"amp-updates/s (QFT, {n}q {'complex128' if prec == capi.C128 else 'complex64'} resident state)",
        "value": upd_step * args.steps / (t1 - t0),
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == capi.C128 else "f32",
        "data": "synthetic: seeded random normalised state generated on the device",
        "config": {"workload": f"QFT-{n} {args.precision} (QFT.cpp:9-22), {len(ops)} gates, state resident in HBM",
                   "qubits": n, "precision": args.precision, "fused": fused, "passes_per_step": passes,
                   "tile_qubits": plan["tile_bits"], "contiguous_low_qubits": plan["low_bits"],
                   "jit": args.jit, "jit_timed": jit_timed, "jit_compile_ms": jit1["compile_ms"],
                   "autotune": tune},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": bytes_per_launch, "mean_launch_ms": mean_ms,
                     "copy_probe_gbs": {"contiguous_64KiB_tiles": copy_contig, "runs_of_the_plan": copy_runs},
                     "ratio_to_copy_probe": achieved / 1e9 / copy_runs if copy_runs > 0 else None},
        "pct_hbm_step": (bytes_per_launch * dom["launches"] / args.steps) / (ms_step * 1e-3) / HBM_PEAK,
    }
    if args.verify > 0:
        # one QFT of the seeded state; QFT^k is not QFT for the timed k steps
        st.set_random(12345)
        st.run(ops, fused=fused)
        rng = np.random.default_rng(n)
        ks = np.concatenate([np.array([0, (1 << n) - 1], dtype=np.uint64),
                             rng.integers(0, 1 << n, size=max(args.verify - 2, 0), dtype=np.uint64)])[:args.verify]
        got = np.array([st.download_range(int(k), 1)[0] for k in ks])
        t0 = time.perf_counter()
        want = ctx.qft_probe(n, 12345, ks)
        log(f"closed-form probe of {len(ks)} amplitudes: {time.perf_counter() - t0:.2f} s")
        err = np.abs(got - want)
        result["verify"] = {"amplitudes": int(len(ks)), "max_abs_err": float(err.max()),
                            "max_rel_err": float(err.max() / np.abs(want).max()),
                            "reference": "FP64 DFT closed form from the seed (qse_qft_probe), one QFT of the "
                                         "seeded random state"}
    st.close()

    return result


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--unfused", action="store_true", help="one sweep per gate instead of fused passes")
    ap.add_argument("--autotune", action="store_true",
                    help="time the tuner's best 3 tile shapes on the state first and keep the fastest")
    ap.add_argument("--jit", choices=["auto", "off", "always"], default="auto",
                    help="run-time specialised passes: compiled on the program's second run (auto), never, or at once")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3, help="timed runs of the CPU baseline (median reported)")
    ap.add_argument("--cpu-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--no-qaoa", action="store_true")
    ap.add_argument("--qaoa-batch", type=int, default=64)
    ap.add_argument("--qaoa-shift-sets", type=int, default=1)
    ap.add_argument("--no-extra", action="store_true", help="skip the noise-trajectory and uneven-batch legs")
    ap.add_argument("--no-dense", action="store_true", help="skip the matrix-core dense-block leg")
    ap.add_argument("--verify", type=int, default=16, metavar="K",
                    help="after timing, rerun one QFT from the seed and compare K output amplitudes with the FP64 "
                         "DFT closed form evaluated on the GPU (qse_qft_probe)")
    ap.add_argument("--legs", default="qft,layered,dense,exchange,qaoa,noise,batch,cpu",
                    help="comma list of parts to run (qft = the headline QFT line; the others are the extra legs); "
                         "a leg alone, e.g. --legs noise, is what a profiler run of that leg uses")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run X on qubit 0 (one read + one write of every amplitude) for counter calibration")
    args = ap.parse_args()
    if args.cpu_child:
        return cpu_child(args.cpu_child, args.cpu_reps)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the library drives one GPU per process and this benchmark measures one (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    n = args.qubits
    prec = capi.C128 if args.precision == "c128" else capi.C64
    b = 16 if prec == capi.C128 else 8
    if args.calibrate_copy:
        ctx = capi.Context(0)
        st = ctx.state(n, prec)
        for _ in range(args.steps):
            st.x(0)
        ctx.sync()
        log(f"copy calibration: {args.steps} x k_pairs X(0) over {n} qubits, {2 * b * 2**n} bytes each")
        return 0
    ops = C.qft(n).ops()
    upd_step = C.amp_updates(ops, n)

    ctx = capi.Context(0)
    ctx.set_jit({"auto": capi.JIT_AUTO, "off": capi.JIT_OFF, "always": capi.JIT_ALWAYS}[args.jit])
    info = capi.device_info(0)
    log(f"device: {info['name']} CUs={info['compute_units']} HBM={info['total_bytes'] / 2**30:.1f} GiB")
    legs = set(args.legs.split(","))
    for flag, names in ((args.no_extra, ("layered", "noise", "batch", "exchange")), (args.no_dense, ("dense",)),
                        (args.no_qaoa, ("qaoa",)), (args.no_cpu, ("cpu",))):
        if flag:
            legs -= set(names)
    result = qft_leg(ctx, args, n, prec, ops, upd_step) if "qft" in legs else {"legs": sorted(legs)}

    if "layered" in legs and n == 30 and prec == capi.C128:
        result["layered30"] = layered_leg(ctx)
    if "dense" in legs:
        result["dense_block"] = dense_block_leg(
            ctx, {"auto": capi.JIT_AUTO, "off": capi.JIT_OFF, "always": capi.JIT_ALWAYS}[args.jit])
    if "exchange" in legs and n == 30 and prec == capi.C128:
        result["pair_exchange"] = exchange_leg(ctx)
    if "qaoa" in legs:
        result["qaoa"] = qaoa_leg(ctx, args.qaoa_batch, args.qaoa_shift_sets)
    if "noise" in legs:
        result["noise_trajectories"] = noise_leg(ctx)
        log(f"noise trajectories: {result['noise_trajectories']['seconds']:.3f} s deferred, "
            f"{result['noise_trajectories']['interleaved']['seconds']:.3f} s interleaved")
    if "batch" in legs:
        result["uneven_batch"] = uneven_batch_leg(ctx)
        log(f"uneven batch: {result['uneven_batch']['seconds']:.3f} s")
    if "cpu" in legs:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(n, ops, args.cpu_reps)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.3e} amp-updates/s")
    print(json.dumps(result), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
