"""Benchmark of the state-vector engine on one MI355X.

Default workload (one step = one full application of the circuit to the
state resident in HBM): the 30-qubit FP64 random circuit, 20 layers of U3 on
every qubit + brick-wall CZ (890 gates, 16 GiB state).  Prints ONE JSON line:
value = gates/s; roofline = the tiled sweep kernel's algorithmic bytes
(one read + one write of the state, 32 * 2^n B per launch) over its average
launch duration, timed with HIP events on the engine's stream; cpu_baseline =
the CPU oracle (oracle/, OpenMP) on a bounded prefix of the same circuit.

  python bench.py [--steps K] [--warmup W] [--quick] [--no-cpu]

The "extra" object reports the other workloads: the 33-qubit QFT circuit
(gates/s, sweep GB/s), the adjoint gradient of the 20-qubit 320-parameter
ansatz (time vs one forward pass and vs GPU parameter shift), 512 batched
ansatz gradients vs a per-circuit loop, and the 30-qubit QIR gate-set
program with multi-controlled gates and 16 mid-circuit measurements
(gates/s, sweep and measurement GB/s, and the CPU oracle on a prefix of it),
the largest QASMBench circuits through the OpenQASM front end, and the u3
and QFT circuits again with sweeps compiled into straight-line kernels at
upload (SVMI_OPT_SPECIALIZE; compile time reported beside the run time),
and the density-matrix engine on 15 qubits (a 30-qubit state).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "sv-sim_amd"))

import numpy as np  # noqa: E402

METRIC = "gates/s + HBM TB/s (% roofline) at 30–33q FP64; adjoint grad time vs param-shift"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on (the CPU baselines use all of them)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_u3_brickwall(n, steps, warmup):
    import svmi
    from svmi import workloads as W
    gates = W.u3_brickwall(n=n, layers=20)
    e = svmi.Engine(n)
    e.upload(gates)
    for _ in range(warmup):
        e.run()
    e.set_option(svmi.OPT_PROFILE, 1)
    sweep_ms, sweep_n, other_ms = 0.0, 0, 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        e.run()
        st = e.stats()
        sweep_ms += st["sweep_ms"]
        sweep_n += st["sweep_launches"]
        other_ms += st["other_ms"]
    dt = time.perf_counter() - t0
    e.set_option(svmi.OPT_PROFILE, 0)
    # untimed: sampling throughput with fed draws
    _, u = W.draws(4096, seed=2002)
    t1 = time.perf_counter()
    _, total = e.measure(u)
    ts = time.perf_counter() - t1
    passes = e.stats()["sweeps"]
    e.close()
    # FP64 multiply-adds per amplitude of the planned program (planner summary)
    _, desc = svmi.plan(n, gates)
    fma = float(desc.split("fma_per_amp=")[1].split()[0])
    return {"gates": len(gates), "dt": dt, "sweep_ms": sweep_ms, "sweep_n": sweep_n, "other_ms": other_ms,
            "passes": passes, "sample_s": ts, "total_prob": total, "fma_per_amp": fma}


def bench_qft(n, steps, seed=2033):
    """QFT-structured circuit on a seeded random input state (on device)."""
    import svmi
    from svmi import workloads as W
    gates = W.qft(n=n)
    e = svmi.Engine(n)
    e.set_random_state(seed)
    e.upload(gates)
    e.run()
    e.set_option(svmi.OPT_PROFILE, 1)
    t0 = time.perf_counter()
    sweep_ms, sweep_n = 0.0, 0
    for _ in range(steps):
        e.run()
        st = e.stats()
        sweep_ms += st["sweep_ms"]
        sweep_n += st["sweep_launches"]
    dt = (time.perf_counter() - t0) / steps
    passes = e.stats()["sweeps"]
    e.close()
    avg = sweep_ms / max(sweep_n, 1) / 1e3
    return {"qubits": n, "gates": len(gates), "ms_per_circuit": dt * 1e3, "gates_per_s": len(gates) / dt,
            "passes": passes, "sweep_GBps": 32.0 * (1 << n) / avg / 1e9}


def bench_specialized(steps=2):
    """The u3 brick-wall (30 qubits) and QFT (33 qubits) circuits with
    SVMI_OPT_SPECIALIZE: sweeps compiled into straight-line kernels with
    hipRTC at upload.  Reports the upload (compile) time separately from the
    per-circuit time (parity with the interpreted sweeps: tests)."""
    import svmi
    from svmi import workloads as W
    out = {}
    for name, n, gates in (("u3_brickwall_n30", 30, W.u3_brickwall(n=30, layers=20)), ("qft_n33", 33, W.qft(n=33))):
        e = svmi.Engine(n)
        e.set_random_state(2033)
        e.set_option(svmi.OPT_SPECIALIZE, 1)
        log(f"[bench] specialised {name}: compiling")
        t0 = time.perf_counter()
        e.upload(gates)
        t_up = time.perf_counter() - t0
        st = e.stats()
        e.run()
        e.set_option(svmi.OPT_PROFILE, 1)
        sweep_ms, sweep_n, walls = 0.0, 0, []
        for _ in range(steps):
            t0 = time.perf_counter()
            e.run()
            walls.append(time.perf_counter() - t0)
            s2 = e.stats()
            sweep_ms += s2["sweep_ms"]
            sweep_n += s2["sweep_launches"]
        e.set_option(svmi.OPT_PROFILE, 0)
        e.close()
        dt = float(np.median(walls))
        avg = sweep_ms / max(sweep_n, 1) / 1e3
        out[name] = {"qubits": n, "gates": len(gates), "upload_s": t_up, "compile_s": st["compile_ms"] / 1e3,
                     "specialized_sweeps": st["specialized_sweeps"], "ms_per_circuit": dt * 1e3,
                     "gates_per_s": len(gates) / dt, "sweep_GBps": 32.0 * (1 << n) / avg / 1e9}
        log(f"[bench] {out[name]}")
    return out


def bench_adjoint(n=20, layers=8, reps=5):
    """<H> and dE/dtheta of the hardware-efficient ansatz (320 parameters)
    by the adjoint sweep, against one forward pass and against GPU
    parameter shift (2 forward passes + <H> per parameter)."""
    import svmi
    from svmi import workloads as W
    gates, slots, P, terms = W.hea_ising(n=n, layers=layers)
    e = svmi.Engine(n)
    e.upload(gates)
    fw, adj = [], []
    for _ in range(reps):
        e.reset_state()
        t0 = time.perf_counter()
        e.run()
        fw.append(time.perf_counter() - t0)
    for _ in range(reps):
        e.reset_state()
        t0 = time.perf_counter()
        E, g = e.adjoint(terms, slots, P)
        adj.append(time.perf_counter() - t0)
    names = ("theta", "phi", "lambda")
    t0 = time.perf_counter()
    ps = np.zeros(P)
    for gi in range(len(gates)):
        for s_ in range(3):
            p = slots[gi, s_]
            if p < 0:
                continue
            v = []
            for sh in (np.pi / 2, -np.pi / 2):
                G2 = gates.copy()
                G2[gi][names[s_]] += sh
                e.reset_state()
                e.upload(G2)
                e.run()
                v.append(e.expval(terms))
            ps[p] = 0.5 * (v[0] - v[1])
    t_ps = time.perf_counter() - t0
    e.close()
    f, a = float(np.median(fw)), float(np.median(adj))
    return {"qubits": n, "params": P, "gates": len(gates), "forward_ms": f * 1e3, "adjoint_ms": a * 1e3,
            "adjoint_over_forward": a / f, "param_shift_ms": t_ps * 1e3, "param_shift_over_adjoint": t_ps / a,
            "max_abs_diff_vs_param_shift": float(np.abs(g - ps).max())}


def bench_batch(count=512, reps=3):
    """Expectation + adjoint gradient of many uneven ansatz circuits: one
    batched call against a per-circuit loop on the same GPU.  One untimed
    warmup of each, then the median of `reps` timed runs of each."""
    import svmi
    from svmi import workloads as W
    items = W.batch_ising(count=count)

    def loop():
        ev2 = np.zeros(len(items))
        gs = []
        for i, it in enumerate(items):
            with svmi.Engine(it["n"]) as e:
                e.upload(it["gates"])
                ev2[i], g = e.adjoint(it["terms"], it["param_slots"], it["n_params"])
                gs.append(g)
        return ev2, gs

    ev, grads = svmi.batch_adjoint(items)
    ev2, g2 = loop()
    tb, tl = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        ev, grads = svmi.batch_adjoint(items)
        tb.append(time.perf_counter() - t0)
    for _ in range(reps):
        t0 = time.perf_counter()
        ev2, g2 = loop()
        tl.append(time.perf_counter() - t0)
    diff = max(float(np.abs(a - b).max()) for a, b in zip(grads, g2))
    t_batch, t_loop = float(np.median(tb)), float(np.median(tl))
    # share of the loop's device work in the largest circuits (work ~ gates * 2^n)
    w = np.array([len(it["gates"]) * float(1 << it["n"]) for it in items])
    big = np.array([it["n"] >= 19 for it in items])
    return {"circuits": len(items), "batched_ms": t_batch * 1e3, "loop_ms": t_loop * 1e3,
            "batched_runs_ms": [t * 1e3 for t in tb], "loop_runs_ms": [t * 1e3 for t in tl],
            "speedup": t_loop / t_batch, "max_abs_diff": max(diff, float(np.abs(ev - ev2).max())),
            "work_share_n_ge_19": float(w[big].sum() / w.sum())}


def bench_fp64():
    """FP64 vector-FMA and matrix-core (MFMA 16x16x4) rates at 1, 2 and 4
    waves per SIMD, eight independent accumulator chains each: the evidence
    for applying fused dense blocks with v_fma_f64 rather than MFMA."""
    import svmi
    out = {}
    for w in (1, 2, 4):
        v, m = svmi.fp64_rates(w)
        out[f"waves_per_simd_{w}"] = {"vfma_tflops": v, "mfma_tflops": m}
        log(f"[bench] fp64 rates, {w} waves/SIMD: v_fma_f64 {v:.1f} TFLOP/s, v_mfma_f64_16x16x4 {m:.1f} TFLOP/s")
    return out


def bench_c1(reps=10):
    """The 12-qubit random-layer circuit (720 gates) from |0..0> plus 1024
    shots: GPU wall time per run (median of reps, each from a reset state,
    synchronised), and the reference CPU engine (oracle/_ref, compiled from
    svsim/src/svsim_cpu_omp.hpp) timing only its sim() as the reference does,
    median of 10 runs on every host core; the oracle restatement when the
    reference build is absent."""
    import svmi
    from oracle import oracle as O
    from svmi import workloads as W
    n = 12
    gates = W.random_layers(n=n, layers=40)
    _, u = W.draws(1024, seed=2001)
    e = svmi.Engine(n)
    e.upload(gates)
    e.run()
    runs, shots_t = [], []
    for _ in range(reps):
        e.reset_state()
        t0 = time.perf_counter()
        e.run()
        runs.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        shots, _ = e.measure(u)
        shots_t.append(time.perf_counter() - t0)
    re, im = e.get_state()
    dev = e.stats()["device_ms"]
    e.close()
    t = float(np.median(runs))
    out = {"qubits": n, "gates": len(gates), "gpu_run_ms": t * 1e3, "gpu_device_ms": dev,
           "gpu_gates_per_s": len(gates) / t, "gpu_1024_shots_ms": float(np.median(shots_t)) * 1e3}
    cores = host_cores()
    if O.ref_available():
        ms = O.ref_sim_times(n, gates, 10, nthreads=cores)
        ore, oim = O.ref_run(n, gates, nthreads=cores)
        kind = "reference"
    else:
        ms = []
        for _ in range(10):
            t0 = time.perf_counter()
            ore, oim = O.run(n, gates, nthreads=cores)
            ms.append((time.perf_counter() - t0) * 1e3)
        kind = "port"
    cpu_ms = float(np.median(ms))
    out["cpu"] = {"value": len(gates) / (cpu_ms * 1e-3), "unit": "gates/s", "cores": cores, "kind": kind,
                  "sim_ms_median_of_10": cpu_ms, "cpu_model": host_cpu_model(),
                  "sample": "whole circuit, sim() only, median of 10 runs"}
    out["max_abs_diff_vs_cpu"] = float(max(np.abs(re - ore).max(), np.abs(im - oim).max()))
    out["shots_equal_oracle"] = bool(np.array_equal(shots, O.measure(n, ore, oim, u)))
    return out


def cpu_param_shift_c4(budget_s=15.0, threads=None):
    """CPU parameter shift (oracle, OpenMP) for the 20-qubit 320-parameter
    ansatz: each parameter needs two forward passes + <H>; a bounded number
    of parameters is timed and the full gradient's time extrapolated from
    them (all 640 passes would take minutes)."""
    from oracle import oracle as O
    from svmi import workloads as W
    threads = threads or host_cores()
    gates, slots, P, terms = W.hea_ising(n=20, layers=8)
    names = ("theta", "phi", "lambda")
    bound = [(gi, s_) for gi in range(len(gates)) for s_ in range(3) if slots[gi, s_] >= 0]
    done, t = 0, 0.0
    for gi, s_ in bound:
        if t >= budget_s:
            break
        t0 = time.perf_counter()
        for sh in (np.pi / 2, -np.pi / 2):
            g = gates.copy()
            g[gi][names[s_]] += sh
            re, im = O.run(20, g, nthreads=threads)
            O.expval(20, re, im, terms, nthreads=threads)
        t += time.perf_counter() - t0
        done += 1
    per = t / done
    return {"value": per * len(bound) * 1e3, "unit": "ms (full 320-parameter gradient, extrapolated)",
            "cores": threads, "kind": "port", "cpu_model": host_cpu_model(),
            "sample": f"{done} of {len(bound)} parameters ({2 * done} forward passes + <H>), {t:.1f} s"}


def c5_cpu_subset(threads=None, reps_grad=2):
    """16 circuits of the 512-circuit batch spanning n = 10..22: <H> on the
    CPU (oracle) against the GPU batch, and the gradient entries of the first
    `reps_grad` parameters of each by CPU parameter shift; CPU time per
    circuit for <H> and, extrapolated from the shifted passes, for its whole
    gradient."""
    import svmi
    from oracle import oracle as O
    from svmi import workloads as W
    threads = threads or host_cores()
    items = W.batch_ising(count=512)
    pick = []
    for n in range(10, 23):
        idx = [i for i, it in enumerate(items) if it["n"] == n]
        if idx:
            pick.append(idx[0])
    for i, it in enumerate(items):
        if len(pick) >= 16:
            break
        if i not in pick and it["n"] >= 20:
            pick.append(i)
    sub = [items[i] for i in sorted(pick)]
    ev, grads = svmi.batch_adjoint(sub)
    names = ("theta", "phi", "lambda")
    t_ev, t_grad_est, dev, gdev = 0.0, 0.0, 0.0, 0.0
    for k, it in enumerate(sub):
        n = it["n"]
        t0 = time.perf_counter()
        re, im = O.run(n, it["gates"], nthreads=threads)
        e_cpu = O.expval(n, re, im, it["terms"], nthreads=threads)
        t1 = time.perf_counter() - t0
        t_ev += t1
        dev = max(dev, abs(e_cpu - ev[k]))
        slots = np.asarray(it["param_slots"]).reshape(len(it["gates"]), 3)
        bound = [(gi, s_) for gi in range(len(it["gates"])) for s_ in range(3) if slots[gi, s_] >= 0]
        for gi, s_ in bound[:reps_grad]:
            vals = []
            for sh in (np.pi / 2, -np.pi / 2):
                g = it["gates"].copy()
                g[gi][names[s_]] += sh
                r2, i2 = O.run(n, g, nthreads=threads)
                vals.append(O.expval(n, r2, i2, it["terms"], nthreads=threads))
            gdev = max(gdev, abs(0.5 * (vals[0] - vals[1]) - grads[k][slots[gi, s_]]))
        t_grad_est += t1 * 2 * len(bound)
    return {"circuits": len(sub), "qubits": sorted(int(it["n"]) for it in sub),
            "cpu_expval_ms": t_ev * 1e3, "cpu_param_shift_ms_extrapolated": t_grad_est * 1e3,
            "max_abs_diff_expval": float(dev), "max_abs_diff_grad_sampled": float(gdev),
            "cores": threads, "kind": "port", "cpu_model": host_cpu_model()}


def bench_qir(n=30, reps=3):
    """QIR gate-set program (svmi_qir_run): 1200 records incl. 1-3-control
    and (n-1)-control gates and 16 mid-circuit measurements, from |0..0>.
    Wall time per program (planning, upload, device work, one sync) and,
    with HIP events per step, the average sweep and measurement durations."""
    import svmi
    from svmi import workloads as W
    g = W.qir_mixed(n=n)
    nm = int(np.count_nonzero(g["op"] == svmi.QIR_OP["Measure"]))
    e = svmi.Engine(n)
    e.qir_run(n, g)
    walls, sweep_ms, sweep_n, meas_ms, meas_n, dev = [], 0.0, 0, 0.0, 0, []
    for prof in (0, 1):
        e.set_option(svmi.OPT_PROFILE, prof)
        for _ in range(reps):
            e.reset_state()
            t0 = time.perf_counter()
            _, outs, _ = e.qir_run(n, g)
            t = time.perf_counter() - t0
            st = e.stats()
            if prof:
                sweep_ms += st["sweep_ms"]
                sweep_n += st["sweep_launches"]
                meas_ms += st["other_ms"]
                meas_n += st["other_launches"]
            else:
                walls.append(t)
                dev.append(st["device_ms"])
    e.close()
    w = float(np.median(walls))
    avg_sweep = sweep_ms / max(sweep_n, 1) / 1e3
    avg_meas = meas_ms / max(meas_n, 1) / 1e3
    return {"qubits": n, "records": len(g), "measures": nm, "ms_per_program": w * 1e3,
            "device_ms": float(np.median(dev)), "gates_per_s": len(g) / w, "sweeps": sweep_n // reps,
            "sweep_ms": avg_sweep * 1e3, "sweep_GBps": 32.0 * (1 << n) / avg_sweep / 1e9,
            "measure_ms": avg_meas * 1e3, "measure_GBps": 16.0 * (1 << (n - 1)) / avg_meas / 1e9,
            "ones": int(outs.sum())}


QASM_CIRCUITS = ("qft_n20", "square_root_n18", "multiplier_n25", "swap_test_n25", "ising_n26", "wstate_n27")


def bench_dm(n=15, reps=3):
    """Density-matrix engine (svmi_dm_qir_run): the QIR gate-set program on
    rho of n = 15 qubits (a 30-qubit, 16 GiB state), 600 records with 8
    mid-circuit measurements, from |0..0><0..0|.  Wall time per program and,
    with HIP events per step, the average sweep and measurement durations."""
    import svmi
    from svmi import workloads as W
    g = W.qir_mixed(n=n, n_gates=600, n_measures=8)
    e = svmi.Engine(2 * n)
    e.dm_qir_run(n, g)
    walls, sweep_ms, sweep_n, meas_ms, meas_n = [], 0.0, 0, 0.0, 0
    for prof in (0, 1):
        e.set_option(svmi.OPT_PROFILE, prof)
        for _ in range(reps):
            e.reset_state()
            t0 = time.perf_counter()
            probs, outs = e.dm_qir_run(n, g)
            t = time.perf_counter() - t0
            st = e.stats()
            if prof:
                sweep_ms += st["sweep_ms"]
                sweep_n += st["sweep_launches"]
                meas_ms += st["other_ms"]
                meas_n += st["other_launches"]
            else:
                walls.append(t)
    e.close()
    w = float(np.median(walls))
    avg_sweep = sweep_ms / max(sweep_n, 1) / 1e3
    return {"qubits": n, "state_qubits": 2 * n, "records": len(g), "measures": len(probs),
            "ms_per_program": w * 1e3, "records_per_s": len(g) / w, "sweeps": sweep_n // reps,
            "sweep_ms": avg_sweep * 1e3, "sweep_GBps": 32.0 * (1 << (2 * n)) / avg_sweep / 1e9,
            "measure_ms": meas_ms / max(meas_n, 1), "ones": int(outs.sum())}


def cpu_baseline_dm(n=15, budget_s=20.0, threads=None):
    """Oracle restatement of the reference density-matrix engine (bit-exact
    to it) on the first records of the same program: one timed call on a
    prefix sized from a one-record calibration to ~budget_s.  (The reference
    engine itself stops at 13 qubits: 2n < N_QUBIT_SLOT = 28.)"""
    from oracle import oracle as O
    from svmi import workloads as W
    threads = threads or host_cores()
    g = W.qir_mixed(n=n, n_gates=600, n_measures=8)
    t0 = time.perf_counter()
    O.dm_qir_apply(n, g[:1], nthreads=threads)
    t1 = time.perf_counter() - t0
    k = int(max(2, min(len(g), budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    O.dm_qir_apply(n, g[:k], nthreads=threads)
    t = time.perf_counter() - t0
    return {"value": k / t, "unit": "records/s", "cores": threads, "kind": "port",
            "sample": f"first {k} of {len(g)} records of the {n}-qubit density-matrix program, {t:.1f} s"}


def bench_qasm(reps=3):
    """The reference's largest QASMBench circuits (tests/golden/qasm/):
    OpenQASM parse time on the host, then median device time per run."""
    import svmi
    from svmi import qasm
    out = {}
    for name in QASM_CIRCUITS:
        with open(os.path.join(ROOT, "tests", "golden", "qasm", name + ".qasm"), "rb") as f:
            text = f.read()
        t0 = time.perf_counter()
        n, gates, info = qasm.parse(text)
        t_parse = time.perf_counter() - t0
        e = svmi.Engine(n)
        t0 = time.perf_counter()
        e.upload(gates)
        t_plan = time.perf_counter() - t0
        runs = []
        for _ in range(reps):
            e.reset_state()
            t0 = time.perf_counter()
            e.run()
            runs.append(time.perf_counter() - t0)
        st = e.stats()
        e.close()
        t = float(np.median(runs))
        out[name] = {"qubits": n, "gates": len(gates), "parse_ms": t_parse * 1e3, "plan_ms": t_plan * 1e3,
                     "run_ms": t * 1e3, "gates_per_s": len(gates) / t, "sweeps": st["sweeps"]}
    return out


def cpu_reference_qasm(name="qft_n20", threads=None):
    """The reference CPU engine itself (oracle/_ref, compiled from
    svsim/src/svsim_cpu_omp.hpp) on one whole QASMBench circuit; None when it
    was not built."""
    from oracle import oracle as O
    from svmi import qasm
    if not O.ref_available():
        return None
    threads = threads or host_cores()
    n, gates, _ = qasm.load(os.path.join(ROOT, "tests", "golden", "qasm", name + ".qasm"))
    t0 = time.perf_counter()
    O.ref_run(n, gates, nthreads=threads)
    t = time.perf_counter() - t0
    return {"value": len(gates) / t, "unit": "gates/s", "cores": threads, "kind": "reference",
            "sample": f"{name} ({n} qubits, {len(gates)} gates), whole circuit, {t:.1f} s"}


def cpu_baseline_qir(n, budget_s=10.0, threads=None):
    """Oracle restatement of the reference QIR engine (bit-identical to it,
    OpenMP) on the first records of the same program, until ~budget_s."""
    from oracle import oracle as O
    from svmi import workloads as W
    threads = threads or host_cores()
    g = W.qir_mixed(n=n)
    re, im = O.zero_state(n)
    done, t = 0, 0.0
    while t < budget_s and done < len(g):
        chunk = g[done:done + 4]
        t0 = time.perf_counter()
        O.qir_apply(n, re, im, chunk, nthreads=threads)
        t += time.perf_counter() - t0
        done += len(chunk)
    return {"value": done / t, "unit": "records/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(g)} records of the {n}-qubit QIR program, {t:.1f} s"}


def measured_traffic(n):
    """HBM bytes per sweep launch of the default circuit, from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (scaled by a copy of known size);
    None for other sizes."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    key = f"u3_brickwall n={n} layers=20"
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(key)
    return None if rec is None else rec["bytes_per_launch"]


def stream_copy_gbps(gib=4.0, reps=5):
    """Measured device copy bandwidth (read + write bytes / time) of a
    `gib`-GiB FP64 buffer, HIP events on the current stream, best of reps."""
    import torch
    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    best = 0.0
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = max(best, 2 * n * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del src, dst
    torch.cuda.empty_cache()
    return best


def calibration_copy(gib):
    """One device-to-device copy of `gib` GiB (a plain copy kernel of known
    traffic: the counters of HBM reads and writes scale by it)."""
    import torch
    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] calibration copy of {n * 8} bytes done")
    del src, dst
    torch.cuda.empty_cache()


def cpu_baseline_u3(n, budget_s=12.0, threads=None):
    """Oracle (OpenMP C restatement of the reference CPU engine) on the first
    gates of the same circuit, in chunks until ~budget_s of CPU work."""
    from oracle import oracle as O
    from svmi import workloads as W
    threads = threads or host_cores()
    gates = W.u3_brickwall(n=n, layers=20)
    re, im = O.zero_state(n)
    done, t = 0, 0.0
    while t < budget_s and done < len(gates):
        chunk = gates[done:done + 5]
        t0 = time.perf_counter()
        rc = O.lib().orc_apply_gates(n, O._ptr(re), O._ptr(im), O._ptr(np.ascontiguousarray(chunk)), len(chunk), threads)
        t += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("oracle failed")
        done += len(chunk)
    return {"value": done / t, "unit": "gates/s", "cores": threads, "kind": "port", "cpu_model": host_cpu_model(),
            "sample": f"first {done} of 890 gates of the {n}-qubit u3 brick-wall circuit, {t:.1f} s "
                      "(the reference engine stops at 28 qubits: 32-bit IdxType)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the QFT, adjoint and batch workloads")
    ap.add_argument("--only", choices=("qft", "adjoint", "batch", "qir", "qasm", "specialized", "dm", "c1",
                                          "fp64"),
                    help="run only this extra workload and print its result (profiling)")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="first run one device copy of this many GiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the engine runs on a single device)",
              file=sys.stderr)
        sys.exit(2)
    # torch's HIP runtime initialises the device before the engine's
    import torch  # noqa: F401
    if a.copy_gib > 0:
        calibration_copy(a.copy_gib)
    if a.only:
        fn = {"qft": lambda: bench_qft(33, steps=a.steps), "adjoint": bench_adjoint, "batch": bench_batch,
              "qir": lambda: bench_qir(30, reps=a.steps), "qasm": lambda: bench_qasm(reps=a.steps),
              "specialized": lambda: bench_specialized(steps=a.steps),
              "dm": lambda: bench_dm(15, reps=a.steps), "c1": bench_c1, "fp64": bench_fp64}[a.only]
        log(f"[bench] only {a.only}")
        print(json.dumps({a.only: fn()}), flush=True)
        return
    copy_gbps = stream_copy_gbps()
    log(f"[bench] device copy {copy_gbps:.0f} GB/s")
    log(f"[bench] u3 brick-wall n={a.qubits} warmup={a.warmup} steps={a.steps}")
    r = bench_u3_brickwall(a.qubits, a.steps, a.warmup)
    ms = r["dt"] / a.steps * 1e3
    value = r["gates"] / (r["dt"] / a.steps)
    bytes_per_sweep = 32.0 * (1 << a.qubits)
    avg_sweep_s = r["sweep_ms"] / max(r["sweep_n"], 1) / 1e3
    achieved = bytes_per_sweep / avg_sweep_s / 1e9
    log(f"[bench] {ms:.1f} ms/step, {r['passes']} passes/step, sweep avg {avg_sweep_s*1e3:.3f} ms, "
        f"{achieved:.0f} GB/s, sampling 4096 shots {r['sample_s']*1e3:.1f} ms")
    extra = {"stream_copy_GBps": copy_gbps}
    log("[bench] C1: 12-qubit random layers + 1024 shots, GPU vs reference CPU engine")
    extra["c1_random_layers_n12"] = bench_c1()
    log(f"[bench] {extra['c1_random_layers_n12']}")
    if not a.quick:
        log("[bench] qft n=33 on a seeded random input")
        extra["qft_n33"] = bench_qft(33, steps=2)
        log(f"[bench] {extra['qft_n33']}")
        extra["specialized"] = bench_specialized()
        log("[bench] adjoint gradient, 20-qubit ansatz, 320 parameters")
        extra["adjoint_hea_n20"] = bench_adjoint()
        log(f"[bench] {extra['adjoint_hea_n20']}")
        if not a.no_cpu:
            extra["adjoint_hea_n20"]["cpu_param_shift"] = cpu_param_shift_c4()
            log(f"[bench] C4 cpu {extra['adjoint_hea_n20']['cpu_param_shift']}")
        log("[bench] 512 batched ansatz gradients vs per-circuit loop")
        extra["batch_ising_512"] = bench_batch()
        log(f"[bench] {extra['batch_ising_512']}")
        if not a.no_cpu:
            extra["batch_ising_512"]["cpu_subset"] = c5_cpu_subset()
            log(f"[bench] C5 cpu subset {extra['batch_ising_512']['cpu_subset']}")
        log("[bench] QIR gate-set program, 30 qubits, 16 mid-circuit measurements")
        extra["qir_mixed_n30"] = bench_qir(30)
        log(f"[bench] {extra['qir_mixed_n30']}")
        if not a.no_cpu:
            extra["qir_mixed_n30"]["cpu_baseline"] = cpu_baseline_qir(30)
            log(f"[bench] qir cpu {extra['qir_mixed_n30']['cpu_baseline']}")
        log("[bench] density-matrix engine, 15 qubits (30-qubit state), 8 measurements")
        extra["dm_qir_n15"] = bench_dm(15)
        if not a.no_cpu:
            extra["dm_qir_n15"]["cpu_baseline"] = cpu_baseline_dm(15)
        log(f"[bench] {extra['dm_qir_n15']}")
        log("[bench] QASMBench circuits (OpenQASM front end)")
        extra["qasmbench"] = bench_qasm()
        if not a.no_cpu:
            extra["qasmbench"]["cpu_baseline"] = cpu_reference_qasm()
        log(f"[bench] {extra['qasmbench']}")
    cpu = None
    if not a.no_cpu:
        log("[bench] cpu baseline (oracle, OpenMP)")
        cpu = cpu_baseline_u3(a.qubits)
        log(f"[bench] cpu {cpu['value']:.3f} gates/s on {cpu['cores']} threads")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"u3_brickwall n={a.qubits} layers=20 (890 gates, FP64, state in HBM)",
                   "qubits": a.qubits, "gates": r["gates"], "passes_per_step": r["passes"],
                   "sweep_GBps": r["passes"] * bytes_per_sweep / (r["dt"] / a.steps) / 1e9,
                   "sampling_shots_per_s": 4096 / r["sample_s"], "fma_per_amp": r["fma_per_amp"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(a.qubits)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
