"""Benchmark of the statevector hot path on one MI355X.

Headline workload (the metric's `value`): random U3 + CNOT-brickwork
circuit, 30 qubits, depth 40, seed 102 (1780 gates), complex64 state resident
in HBM (8 GiB), tiled multi-gate passes.  One step = one execution of the
whole circuit from |0...0> on the resident state (the first pass generates
the start state instead of reading it; the state is left in the qubit layout
of the last pass, which every later read by index restores or maps).

The default run prints ONE JSON line with, besides the headline:
  * the roofline of the dominant kernel (the tile pass, timed with HIP events
    on its own stream inside the timed region) and the streaming rate of a
    plain device copy (k_copy) of the same state;
  * accuracy: the complex64 and complex128 GPU results of the same circuit
    family at `--acc-qubits` against the FP64 oracle (infidelity, max
    |<Z_i Z_i+1>| error), and complex64 against complex128 at 30 qubits;
  * the FP64 CPU oracle timed on this host (the accuracy run's oracle
    execution of the whole circuit), with cores, CPU model and RAM;
  * "c1": the 20-qubit complex128 parity circuit on the GPU and the CPU;
  * "c3_batch": 4096 batched variational circuits (circuits/s, roofline of
    the persistent launch, per-circuit launches, CPU);
  * "c4_serve": a Poisson request stream (p50/p99, saturation, per-circuit
    launches, CPU);
  * "c5_memory_ceiling": 34-qubit complex64 / 33-qubit complex128 QFT +
    Trotter with the FP64 oracle check of <Z0 Z1>.
`--workload batch|serve|qft34` runs one of those alone (its own JSON line).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(args, tile, compiled):
    """HBM bytes per launch of the dominant kernel, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/traffic.json (scaled by
    the plain-copy calibration) for the plan that ran (tile width, compiled or
    interpreted passes); None when no measurement matches this kernel and
    workload shape."""
    if args.lds or args.low not in (0, 5) or args.qubits != 30 or args.depth != 40:
        return None
    key = f"nqpu_pass_compiled_t{tile}" if compiled else f"k_tile_reg_t{tile}"
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)[key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def host_info():
    """CPU model, logical CPUs and RAM of the host the CPU baselines ran on."""
    model, ram = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    ram = round(int(line.split()[1]) / (1 << 20), 1)
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "ram_gib": ram}


def cpu_baseline(seconds: float, n: int = 26):
    """FP64 oracle (C + OpenMP, all host threads), gate by gate like the
    reference's GateOperations path, on the same circuit family at n qubits
    until `seconds` of work are done."""
    from oracle import cpu_ref
    from nqpu_amd.workloads import random_u3_brickwork
    c = random_u3_brickwork(n, 40, 102)
    a = cpu_ref.zero_state(n)
    done = 0
    t0 = time.perf_counter()
    for g in c.gates:
        cpu_ref.apply_gate(a, n, g)
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": done * float(1 << n) / dt, "unit": "amp-updates/s", "cores": cpu_ref.num_threads(),
            "kind": "port",
            "sample": f"first {done} gates of the same U3+CNOT brickwork family at {n} qubits, complex128, "
                      f"one full-state pass per gate ({dt:.1f} s)"}


def bench_batch(args):
    """4096 hardware-efficient-ansatz circuits (8-20 qubits, 4-12 layers of
    RY.RZ + CZ ladder, seed 103), each returning its transverse-field Ising
    energy: one persistent launch versus plain per-circuit launches on the
    same GPU and the FP64 oracle on the host."""
    import numpy as np

    from nqpu_amd import engine as E
    from nqpu_amd.workloads import variational_batch

    if E.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    t0 = time.perf_counter()
    vcs = variational_batch(args.circuits, seed=103)
    circs = [v.circuit.native() for v in vcs]
    obs = [v.hamiltonian() for v in vcs]
    obs_native = E.Observables(obs)  # the term table, built once like the native circuits
    log(f"built {len(circs)} circuits in {time.perf_counter() - t0:.1f} s "
        f"({sum(len(v.circuit.gates) for v in vcs)} gates)")
    b = E.Batch()
    steps, warmup = args.batch_steps or args.steps, args.batch_warmup if args.batch_warmup >= 0 else args.warmup
    for w in range(warmup):
        b.expectation(circs, obs_native)
        log(f"warmup {w + 1}/{warmup} done")
    # batches are queued without waiting: the host plans batch k+1 while the
    # device runs batch k (two in flight); the timed region ends when the
    # last batch's energies are on the host
    t1 = time.perf_counter()
    outs = [b.expectation(circs, obs_native, async_=True) for _ in range(steps)]
    b.synchronize()
    elapsed = time.perf_counter() - t1
    energies = outs[-1]
    st = b.last_stats()
    value = len(circs) * steps / elapsed
    kms, pms = st["kernel_ms"], st["plan_ms"]
    log(f"batched: {elapsed / steps * 1e3:.1f} ms per batch (last batch: kernel {kms:.1f} ms, host {pms:.1f} ms)")
    achieved = st["hbm_bytes"] / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    # plain per-circuit launches: one state, one nqpu_run and one Pauli sum per circuit
    sample = list(range(0, len(circs), max(1, len(circs) // args.per_circuit_sample)))
    t2 = time.perf_counter()
    per = []
    for k in sample:
        n = vcs[k].circuit.num_qubits
        with E.StateVector(n, "c64") as sv:
            sv.run(circs[k])
            per.append(sv.expectation_pauli_sum(obs[k]))
    per_rate = len(sample) / (time.perf_counter() - t2)
    log(f"per-circuit launches: {per_rate:.1f} circuits/s over {len(sample)} circuits")
    result = {
        "metric": "batched circuits/s (4096 variational circuits, 8-20 qubits)",
        "value": value, "unit": "circuits/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": "4096 hardware-efficient ansatz circuits, n ~ U{8..20}, L ~ U{4..12}, "
                               "RY.RZ layers + CZ ladder, TFIM energy with g ~ U[0.2,1.8], seed 103, "
                               "one persistent launch", "circuits": len(circs)},
        "kernel_ms_per_batch": kms, "plan_ms_per_batch": pms,
        "pipelined": "two batches in flight: host planning of batch k+1 overlaps the launch of batch k",
        "batch_stats": st,
        # the persistent launch: algorithmic state bytes of its HBM-resident
        # circuits (2 x gate passes - 1 + observable passes) x 2^n x 8 over its
        # event-timed duration
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_batch",
                     "avg_launch_ms": kms, "bytes_per_launch": st["hbm_bytes"]},
        "gpu_per_circuit": {"circuits_per_s": per_rate, "sampled_circuits": len(sample)},
        "speedup_vs_per_circuit": value / per_rate,
        "per_circuit_vs_batched_max_abs_diff": float(np.max(np.abs(np.array(per) - energies[sample]))),
    }
    if not args.no_cpu:
        log("cpu baseline")
        from oracle import cpu_ref
        # the oracle's energies of circuits of up to 16 qubits check the batched ones
        check = [k for k in sample if vcs[k].circuit.num_qubits <= 16][:48]
        errs = []
        for k in check:
            c = vcs[k].circuit
            a = cpu_ref.run(c.num_qubits, c.gates)
            ref = sum(cf * cpu_ref.expectation_pauli(a, c.num_qubits, *E.pauli_masks(p)) for cf, p in obs[k])
            errs.append(abs(ref - energies[k]))
        result["energy_max_abs_err_vs_fp64_oracle"] = max(errs) if errs else None
        result["checked_circuits"] = len(check)
        t3 = time.perf_counter()
        done = 0
        for k in sample:
            c = vcs[k].circuit
            a = cpu_ref.run(c.num_qubits, c.gates)
            for cf, p in obs[k]:
                cpu_ref.expectation_pauli(a, c.num_qubits, *E.pauli_masks(p))
            done += 1
            if time.perf_counter() - t3 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t3
        result["cpu_baseline"] = {"value": done / dt, "unit": "circuits/s", "cores": cpu_ref.num_threads(),
                                  "kind": "port",
                                  "sample": f"{done} circuits of the same batch (every {len(circs) // len(sample)}th), "
                                            f"complex128, gate by gate, serial Pauli sums ({dt:.1f} s)"}
        result["speedup_vs_cpu"] = value / result["cpu_baseline"]["value"]
    b.close()
    return result


def bench_qft_trotter(args):
    """Memory-ceiling state: a 34-qubit complex64 state (128 GiB) runs a QFT
    followed by 20 first-order Trotter steps (dt = 0.05, g = 1) of the 1D
    transverse-field Ising chain, then a 33-qubit complex128 state (128 GiB)
    runs the same circuit.  One step = the whole circuit from |0...0> (the
    first pass generates the start state) and (complex64) the FP64
    renormalisation of the result.  <Z0 Z1> after
    this circuit depends only on qubits 0..21 (each Trotter step widens the
    light cone of Z0 Z1 by one qubit and the QFT of
    |0...0> is a product state), so the FP64 oracle run at 22 qubits gives
    the exact value for every size."""
    from nqpu_amd import engine as E
    from nqpu_amd.workloads import qft_trotter

    if E.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    n = args.qubits if args.qubits != 30 else 34
    n128 = n - 1
    steps, warmup = args.qft_steps or args.steps, args.qft_warmup if args.qft_warmup >= 0 else args.warmup
    circ = qft_trotter(n, 20)
    nc = circ.native()
    gates = len(circ.gates)
    run_kw = dict(fusion="tiled", tile_qubits=args.tile, low_qubits=args.low, lds_tiles=args.lds,
                  mfma=not args.no_mfma, compile=not (args.interpret or args.lds))
    log(f"workload n={n} gates={gates} plan={nc.plan('c64', **run_kw)}")
    zz = "ZZ" + "I" * (n - 2)
    sv = E.StateVector(n, "c64")
    for w in range(warmup):
        sv.run(nc, from_zero=True, **run_kw)
        log(f"warmup {w + 1}/{warmup} done")
    pass_ms, timed_passes, passes = 0.0, 0, 0
    t0 = time.perf_counter()
    for k in range(steps):
        st = sv.run(nc, time_passes=True, from_zero=True, **run_kw)
        norm_raw = sv.norm()
        sv.normalize()  # FP64 renormalisation on the GPU
        pass_ms += st["pass_ms"]
        timed_passes += st["timed_passes"]
        passes = st["passes"]
        log(f"step {k + 1}/{steps} done")
    elapsed = time.perf_counter() - t0
    dim = 1 << n
    value = gates * dim * steps / elapsed
    bytes_per_pass = 2.0 * dim * 8
    avg_pass_ms = pass_ms / max(timed_passes, 1)
    achieved = bytes_per_pass / (avg_pass_ms * 1e-3) / 1e9
    z64 = sv.expectation_pauli(zz)
    norm64 = sv.norm()
    sv.close()
    log(f"c64: {elapsed / steps:.3f} s per step, <Z0Z1> = {z64!r}, norm = {norm64!r}")

    circ128 = qft_trotter(n128, 20)
    nc128 = circ128.native()
    sv = E.StateVector(n128, "c128")
    # complex128: compiled passes on 11-qubit tiles (register/lane exchanges);
    # the first run generates and compiles the pass kernels, the second is timed
    kw128 = dict(run_kw, compile=not (args.interpret or args.lds), tile_qubits=args.tile or 11)
    t1 = time.perf_counter()
    sv.run(nc128, from_zero=True, **kw128)
    first128 = time.perf_counter() - t1
    log(f"c128 first run (with kernel compilation): {first128:.1f} s")
    t1 = time.perf_counter()
    st128 = sv.run(nc128, time_passes=True, from_zero=True, **kw128)
    dt128 = time.perf_counter() - t1
    z128 = sv.expectation_pauli("ZZ" + "I" * (n128 - 2))
    norm128 = sv.norm()
    sv.close()
    log(f"c128 {n128} qubits: {dt128:.3f} s, <Z0Z1> = {z128!r}, norm = {norm128!r}")
    result = {
        "metric": f"amp-updates/s ({n}-qubit complex64 QFT + 20 Trotter steps)",
        "value": value, "unit": "amp-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"QFT + 20 first-order TFIM Trotter steps (dt 0.05, g 1), {n} qubits complex64 "
                               f"(128 GiB), {n128} qubits complex128 check", "qubits": n, "gates": gates,
                   "passes_per_step": passes},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_tile" if args.lds else ("nqpu_pass_* (compiled)" if run_kw["compile"] else "k_tile_reg"),
                     "avg_launch_ms": avg_pass_ms,
                     "bytes_per_launch": bytes_per_pass},
        "z0z1_c64": z64, "norm_c64": norm64, "norm_c64_before_renormalisation": norm_raw,
        "c128": {"qubits": n128, "seconds": dt128, "passes": st128["passes"], "z0z1": z128, "norm": norm128,
                 "amp_updates_per_s": len(circ128.gates) * float(1 << n128) / dt128,
                 "compiled": kw128["compile"], "first_run_s": first128},
        "z0z1_abs_diff_c64_vs_c128": abs(z64 - z128),
    }
    if not args.no_cpu:
        log("cpu baseline: FP64 oracle, same circuit at 22 and 26 qubits")
        from oracle import cpu_ref
        a = cpu_ref.run(22, qft_trotter(22, 20).gates)
        zref = cpu_ref.expectation_pauli(a, 22, 0, 3)
        result["z0z1_fp64_oracle_22q"] = zref
        result["z0z1_abs_err_c64_vs_oracle"] = abs(z64 - zref)
        result["z0z1_abs_err_c128_vs_oracle"] = abs(z128 - zref)
        nb = 26
        cb = qft_trotter(nb, 20)
        a = cpu_ref.zero_state(nb)
        done = 0
        t3 = time.perf_counter()
        for g in cb.gates:
            cpu_ref.apply_gate(a, nb, g)
            done += 1
            if time.perf_counter() - t3 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t3
        result["cpu_baseline"] = {"value": done * float(1 << nb) / dt, "unit": "amp-updates/s",
                                  "cores": cpu_ref.num_threads(), "kind": "port",
                                  "sample": f"first {done} gates of the same circuit at {nb} qubits, complex128, "
                                            f"one full-state pass per gate ({dt:.1f} s)"}
        result["speedup_vs_cpu"] = value / result["cpu_baseline"]["value"]
    return result


def _percentile(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p)) if len(xs) else None


def bench_serve(args):
    """Served stream: Poisson arrivals at --qps of requests with n from
    {10,14,18,22,26,28} (weights 30/25/20/12/8/5%), random U3+CNOT depth
    U{10..100}, 1024 shots and 8 random Pauli strings each (seed 104).  The
    request server admits them by slot and HBM arena and runs them
    concurrently; latency is arrival -> result.  Compared with plain
    per-circuit launches of the same requests, one after another, whose
    results must be bit-identical."""
    import threading

    import numpy as np

    from nqpu_amd import engine as E
    from nqpu_amd.workloads import served_stream

    if E.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    count = args.requests
    reqs = served_stream(count, args.qps, seed=104)
    natives = [r.circuit.native() for r in reqs]
    gates = sum(len(r.circuit.gates) for r in reqs)
    log(f"{count} requests, {gates} gates, arrivals over {reqs[-1].arrival:.1f} s at {args.qps}/s")
    srv = E.Server(slots=args.slots, precision="c64", memory_bytes=64 << 30)
    for r, nc in list(zip(reqs, natives))[: max(1, args.warmup) * 4]:
        srv.result(srv.submit(nc, r.terms, r.shots, r.seed))
    log("warmup done")
    ids = [None] * count
    lag = [0.0] * count

    def submitter(t0):
        for k, (r, nc) in enumerate(zip(reqs, natives)):
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            lag[k] = max(0.0, time.perf_counter() - t0 - r.arrival)
            ids[k] = srv.submit(nc, r.terms, r.shots, r.seed)

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter, args=(t0,))
    th.start()
    results = []
    for k in range(count):
        while ids[k] is None:
            time.sleep(0.0005)
        results.append(srv.result(ids[k], timeout=600))
        if (k + 1) % 100 == 0:
            log(f"{k + 1}/{count} served")
    th.join()
    elapsed = time.perf_counter() - t0
    lat = [1e3 * lag[k] + results[k]["latency_ms"] for k in range(count)]
    st = srv.stats()
    # saturation: the same requests all at once
    t1 = time.perf_counter()
    burst = [srv.submit(nc, r.terms, r.shots, r.seed) for r, nc in zip(reqs, natives)]
    for i in burst:
        srv.result(i, timeout=600)
    sat = count / (time.perf_counter() - t1)
    srv.close()
    log(f"served: p50 {_percentile(lat, 50):.1f} ms, p99 {_percentile(lat, 99):.1f} ms; burst {sat:.1f}/s")
    # plain per-circuit launches, one after another (own state, run, 8 Pauli strings, shots)
    service = []
    diff = 0.0
    same = True
    for k, (r, nc) in enumerate(zip(reqs, natives)):
        ta = time.perf_counter()
        with E.StateVector(r.circuit.num_qubits, "c64") as sv:
            sv.run(nc)
            ex = np.array([c * sv.expectation_pauli(p) for c, p in r.terms])
            smp = sv.sample(r.shots, seed=r.seed)
        service.append(time.perf_counter() - ta)
        diff = max(diff, float(np.max(np.abs(ex - results[k]["expectations"]))))
        same = same and bool(np.array_equal(smp, results[k]["samples"]))
    seq_rate = count / sum(service)
    fin, seq_lat = 0.0, []
    for r, s in zip(reqs, service):  # FIFO service of the same arrivals, one request at a time
        fin = max(fin, r.arrival) + s
        seq_lat.append(1e3 * (fin - r.arrival))
    result = {
        "metric": f"served circuits/s at {args.qps:g} req/s offered (Poisson), p50/p99 latency",
        "value": count / elapsed, "unit": "circuits/s", "n_gpus": 1, "steps": count, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / count, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": "served stream: Poisson arrivals, n in {10,14,18,22,26,28} (30/25/20/12/8/5%), "
                               "random U3+CNOT depth U{10..100}, 1024 shots + 8 random Pauli strings, seed 104",
                   "requests": count, "offered_qps": args.qps, "slots": args.slots, "gates": gates},
        "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99), "max": max(lat)},
        "saturation_circuits_per_s": sat,
        "server_stats": st,
        "per_circuit_launches": {"circuits_per_s": seq_rate,
                                 "fifo_latency_ms_p50": _percentile(seq_lat, 50),
                                 "fifo_latency_ms_p99": _percentile(seq_lat, 99)},
        "speedup_vs_per_circuit": sat / seq_rate,
        "served_vs_alone_max_abs_diff": diff, "served_vs_alone_samples_identical": same,
    }
    if not args.no_cpu:
        log("cpu baseline")
        from oracle import cpu_ref
        t3 = time.perf_counter()
        done = 0
        for r in reqs:
            n = r.circuit.num_qubits
            a = cpu_ref.run(n, r.circuit.gates)
            for c, p in r.terms:
                cpu_ref.expectation_pauli(a, n, *E.pauli_masks(p))
            cpu_ref.sample_cabi(a, n, r.shots, seed=r.seed)
            done += 1
            if time.perf_counter() - t3 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t3
        result["cpu_baseline"] = {"value": done / dt, "unit": "circuits/s", "cores": cpu_ref.num_threads(),
                                  "kind": "port",
                                  "sample": f"first {done} requests of the same stream served one after another, "
                                            f"complex128 ({dt:.1f} s)"}
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    return result


def bench_reductions(args):
    """Reductions and sampling on the headline state family: 1000 shots,
    all Z_i Z_i+1 (one grouped pass), the transverse-field Ising energy
    (2n-1 terms), all <Z_i>, the norm; best of 5 each, at 30 and 20 qubits
    (complex64) and 30 qubits (complex128)."""
    from nqpu_amd import engine as E
    from nqpu_amd.workloads import random_u3_brickwork

    out = {}
    for n, prec in ((args.qubits, "c64"), (20, "c64"), (args.qubits, "c128")):
        c = random_u3_brickwork(n, 6, 102)
        zz = [(0, 3 << i) for i in range(n - 1)]
        tfim = [(-1.0, "I" * i + "ZZ" + "I" * (n - i - 2)) for i in range(n - 1)] + \
               [(-0.7, "I" * i + "X" + "I" * (n - i - 1)) for i in range(n)]
        with E.StateVector(n, prec) as sv:
            sv.run(c)
            legs = {"sample_1000": lambda: sv.sample(1000, seed=7),
                    "zz_all_grouped": lambda: sv.expectation_pauli_values(zz),
                    "tfim_energy": lambda: sv.expectation_pauli_sum(tfim),
                    "z_all": lambda: sv.expectation_z_all(),
                    "norm": lambda: sv.norm()}
            res = {}
            for name, fn in legs.items():
                fn()
                best = None
                for _ in range(5):
                    t = time.perf_counter()
                    fn()
                    dt = time.perf_counter() - t
                    best = dt if best is None else min(best, dt)
                res[name + "_ms"] = best * 1e3
        out[f"{n}q_{prec}"] = res
        log(f"reductions {n}q {prec}: " + ", ".join(f"{k} {v:.2f}" for k, v in res.items()))
    return out


def bench_c1(args):
    """C1: 20 qubits, 20 layers of U3 on every qubit + CNOT brickwork, seed
    101.  The GPU runs it in complex128 and complex64 (best of 5 after a
    warm-up); the FP64 oracle runs it on the host; the complex128 amplitudes
    and <Z_i> are compared with the oracle's."""
    import numpy as np

    from nqpu_amd import engine as E
    from nqpu_amd.workloads import random_u3_brickwork
    n = 20
    circ = random_u3_brickwork(n, 20, 101)
    nc = circ.native()
    gates = len(circ.gates)
    out = {"workload": "random U3 + CNOT brickwork, 20 qubits, 20 layers, seed 101", "gates": gates}
    amps = {}
    for prec in ("c128", "c64"):
        with E.StateVector(n, prec) as sv:
            sv.run(nc)
            ts = []
            for _ in range(5):
                t = time.perf_counter()
                sv.run(nc, from_zero=True)
                ts.append(time.perf_counter() - t)
            amps[prec] = sv.amplitudes()
            z = sv.expectation_z_all()
            # the same run recorded once as a HIP graph (|0...0> fill + passes) and replayed
            with E.RunGraph(sv, nc, from_zero=True) as g:
                g.launch()
                tg = []
                for _ in range(5):
                    t = time.perf_counter()
                    g.launch()
                    tg.append(time.perf_counter() - t)
                graph_equal = bool(np.array_equal(sv.amplitudes(), amps[prec]))
                ginfo = g.info()
        out[prec] = {"ms": min(ts) * 1e3, "amp_updates_per_s": gates * float(1 << n) / min(ts),
                     "graph_replay_ms": min(tg) * 1e3, "graph_launches": ginfo["launches"],
                     "graph_replay_equals_run": graph_equal}
        amps[prec + "_z"] = z
    if not args.no_cpu:
        from oracle import cpu_ref
        t = time.perf_counter()
        ref = cpu_ref.run(n, circ.gates)
        dt = time.perf_counter() - t
        zref = np.array([cpu_ref.expectation_z(ref, n, q) for q in range(n)])
        out["cpu_baseline"] = {"value": gates * float(1 << n) / dt, "unit": "amp-updates/s",
                               "cores": cpu_ref.num_threads(), "kind": "port", "seconds": dt,
                               "sample": "the whole circuit, complex128, one full-state pass per gate"}
        out["c128_max_abs_amp_err_vs_oracle"] = float(np.max(np.abs(amps["c128"] - ref)))
        out["c128_max_abs_z_err_vs_oracle"] = float(np.max(np.abs(amps["c128_z"] - zref)))
        ov = np.vdot(amps["c64"], ref)
        out["c64_infidelity_vs_oracle"] = float(1 - abs(ov) ** 2 / np.vdot(amps["c64"], amps["c64"]).real)
        out["c64_max_abs_z_err_vs_oracle"] = float(np.max(np.abs(amps["c64_z"] - zref)))
        out["speedup_c128_vs_cpu"] = dt / (out["c128"]["ms"] * 1e-3)
    return out


def bench_variational_loop(args, iters=40):
    """A variational loop on one state: 18-qubit hardware-efficient ansatz
    (RY.RZ on every qubit per layer + CZ ladder, 8 layers), each iteration new
    angles and the transverse-field Ising energy (g = 1).  Plain: nqpu_run per
    iteration (plan, upload, launches).  Graph: the run recorded once as a
    HIP graph; each iteration copies the new angles' register programs into
    it (nqpu_graph_update) and replays it.  complex64, energies in FP64."""
    import numpy as np

    from nqpu_amd import engine as E
    n, layers = 18, 8
    rng = np.random.default_rng(11)

    def ansatz():
        c = E.Circuit(n)
        for _ in range(layers):
            for q in range(n):
                c.ry(q, float(rng.uniform(-np.pi, np.pi)))
                c.rz(q, float(rng.uniform(-np.pi, np.pi)))
            for q in range(n - 1):
                c.cz(q, q + 1)
        return c.native()

    circs = [ansatz() for _ in range(iters)]
    ham = [(-1.0, "I" * i + "ZZ" + "I" * (n - i - 2)) for i in range(n - 1)] + \
          [(-1.0, "I" * i + "X" + "I" * (n - i - 1)) for i in range(n)]
    out = {"workload": f"{n}-qubit ansatz, {layers} layers RY.RZ + CZ ladder, TFIM energy, {iters} iterations",
           "gates": (2 * n + n - 1) * layers}
    with E.StateVector(n, "c64") as sv:
        sv.run(circs[0])
        sv.expectation_pauli_sum(ham)
        t = time.perf_counter()
        e_plain = []
        for c in circs:
            sv.run(c, from_zero=True)
            e_plain.append(sv.expectation_pauli_sum(ham))
        dt_plain = time.perf_counter() - t
        with E.RunGraph(sv, circs[0], from_zero=True) as g:
            g.launch()
            t = time.perf_counter()
            e_graph = []
            for c in circs:
                g.update(c)
                g.launch(async_=True)
                e_graph.append(sv.expectation_pauli_sum(ham))
            dt_graph = time.perf_counter() - t
            info = g.info()
    out["plain_ms_per_iteration"] = dt_plain * 1e3 / iters
    out["graph_ms_per_iteration"] = dt_graph * 1e3 / iters
    out["speedup_graph_vs_plain"] = dt_plain / dt_graph
    out["graph_recaptures"] = info["recaptures"]
    out["graph_plan_ms_last_update"] = info["plan_ms"]
    out["energy_max_abs_diff_graph_vs_plain"] = float(np.max(np.abs(np.array(e_plain) - np.array(e_graph))))
    return out


def accuracy_vs_oracle(args, run_kw):
    """The headline circuit family at --acc-qubits qubits (depth 40, seed
    102) on the FP64 oracle and on the GPU in complex64 and complex128.  The
    oracle state is uploaded into a complex128 GPU state; infidelities are
    FP64 overlaps on the device and the <Z_i Z_i+1> of the oracle state are
    the complex128 reductions over it.  The oracle's execution is timed as
    the CPU baseline (the whole circuit, gate by gate)."""
    from nqpu_amd import engine as E
    from nqpu_amd.workloads import random_u3_brickwork
    from oracle import cpu_ref

    n = args.acc_qubits
    circ = random_u3_brickwork(n, args.depth, 102)
    gates = len(circ.gates)
    log(f"accuracy: FP64 oracle at {n} qubits ({gates} gates, {cpu_ref.num_threads()} threads)")
    t = time.perf_counter()
    ref = cpu_ref.run(n, circ.gates)
    dt = time.perf_counter() - t
    log(f"oracle done in {dt:.1f} s")
    out = {"qubits": n, "depth": args.depth, "gates": gates}
    with E.StateVector(n, "c128") as o:
        o.set_amplitudes(ref)
        del ref
        zz_terms = [(0, 3 << i) for i in range(n - 1)]
        zz_ref = o.expectation_pauli_values(zz_terms)
        for prec in ("c64", "c128"):
            with E.StateVector(n, prec) as sv:
                sv.run(circ.native(), **(run_kw if prec == "c64" else {}))
                ov = o.overlap(sv)
                inf = 1.0 - abs(ov) ** 2 / (sv.norm() * o.norm())
                zz = sv.expectation_pauli_values(zz_terms)
            out[f"infidelity_{prec}_vs_oracle"] = inf
            out[f"zz_max_abs_err_{prec}_vs_oracle"] = float(abs(zz - zz_ref).max())
    cpu = {"value": gates * float(1 << n) / dt, "unit": "amp-updates/s", "cores": cpu_ref.num_threads(),
           "kind": "port",
           "sample": f"the whole U3+CNOT brickwork circuit of the same family at {n} qubits (depth {args.depth}, "
                     f"{gates} gates), complex128, one full-state pass per gate ({dt:.1f} s)"}
    return out, cpu


def copy_probe(n):
    """Streaming rate of a plain device copy (k_copy) of a 2^n complex64
    state: 2 x 2^n x 8 bytes per copy, best of 5."""
    from nqpu_amd import engine as E
    with E.StateVector(n, "c64") as a, E.StateVector(n, "c64") as b:
        b.copy_from(a)
        best = None
        for _ in range(5):
            t = time.perf_counter()
            b.copy_from(a)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
    gbs = 2.0 * (8 << n) / best / 1e9
    return {"kernel": "k_copy", "bytes": 2 * (8 << n), "ms": best * 1e3, "achieved_gbs": gbs,
            "frac_of_peak": gbs / HBM_PEAK_GBS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="random30", choices=["random30", "batch", "qft34", "serve", "reductions"],
                    help="random30: the 30-qubit random circuit headline with every other config as extra "
                         "keys (default); batch / qft34 / serve: one of those alone")
    ap.add_argument("--qps", type=float, default=40.0, help="serve: offered requests per second")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests in the stream")
    ap.add_argument("--slots", type=int, default=8, help="serve: concurrent requests on the GPU")
    ap.add_argument("--circuits", type=int, default=4096)
    ap.add_argument("--per-circuit-sample", type=int, default=512,
                    help="batch: circuits run with plain per-circuit launches for comparison")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-steps", type=int, default=0, help="batch steps (0 = --steps)")
    ap.add_argument("--batch-warmup", type=int, default=-1, help="batch warm-up (-1 = --warmup)")
    ap.add_argument("--qft-steps", type=int, default=0, help="qft34 steps (0 = --steps)")
    ap.add_argument("--qft-warmup", type=int, default=-1, help="qft34 warm-up (-1 = --warmup)")
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--acc-qubits", type=int, default=28,
                    help="qubits of the FP64 oracle accuracy run of the headline circuit family")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--low", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy runs")
    ap.add_argument("--no-per-gate", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner: fixed default plan with compiled complex64 passes")
    ap.add_argument("--only-headline", action="store_true",
                    help="the 30-qubit headline without the other configs (for profiling)")
    ap.add_argument("--lds", action="store_true", help="LDS-resident tile kernel instead of the register one")
    ap.add_argument("--interpret", action="store_true",
                    help="run register tile passes with the interpreting kernel instead of per-pass kernels "
                         "compiled at run time")
    ap.add_argument("--no-mfma", action="store_true",
                    help="keep every gate on the vector ALUs (no fused blocks on the matrix cores)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain state copy (k_copy) of known size, for counter calibration")
    args = ap.parse_args()
    if args.per_circuit_sample < 1:
        ap.error("--per-circuit-sample must be at least 1 (the batch leg compares against that many circuits)")
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the statevector path runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "qft34":
        print(json.dumps(bench_qft_trotter(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.workload == "reductions":
        print(json.dumps(bench_reductions(args)), flush=True)
        return

    from nqpu_amd import engine as E
    from nqpu_amd.workloads import random_u3_brickwork

    if E.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    n, depth = args.qubits, args.depth
    circ = random_u3_brickwork(n, depth, 102)
    nc = circ.native()
    gates = len(circ.gates)
    tuned = None
    if args.no_tune or args.lds or args.interpret or args.tile or args.low or args.no_mfma:
        run_kw = dict(fusion="tiled", tile_qubits=args.tile, low_qubits=args.low, lds_tiles=args.lds,
                      mfma=not args.no_mfma, compile=not (args.interpret or args.lds))
    else:
        # the autotuner times the plan choices on this circuit (before the timed
        # region) and its choice becomes the default plan for 30-qubit complex64
        t0 = time.perf_counter()
        tuned = E.autotune(n, "c64", probe=nc)
        tuned["tuning_s"] = time.perf_counter() - t0
        log(f"autotuned plan: {tuned}")
        run_kw = dict(fusion="tiled")
    plan_tile = tuned["tile_qubits"] if tuned else (args.tile or 12)  # 12: the default for 30 qubits
    plan_compiled = bool(tuned["compile"]) if tuned else bool(run_kw["compile"])
    log(f"workload n={n} depth={depth} gates={gates} plan={nc.plan('c64', **run_kw)}")

    sv = E.StateVector(n, "c64")
    if args.copy_probe:
        tmp = E.StateVector(n, "c64")
        tmp.copy_from(sv)
        tmp.close()
        log(f"copy probe: {2 * (8 << n)} bytes moved by k_copy")
    warm_s = []
    for w in range(args.warmup):
        tw = time.perf_counter()
        sv.run(nc, from_zero=True, **run_kw)
        warm_s.append(time.perf_counter() - tw)
        log(f"warmup {w + 1}/{args.warmup} done ({warm_s[-1]:.2f} s)")
    pass_ms, timed_passes, passes = 0.0, 0, 0
    sv.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        # queued without waiting: the host plans step k+1 while step k runs;
        # each step runs the whole circuit from |0...0>
        st = sv.run(nc, time_passes=True, async_=True, from_zero=True, **run_kw)
        passes = st["passes"]
    st = sv.synchronize()
    elapsed = time.perf_counter() - t0
    pass_ms, timed_passes = st["pass_ms"], st["timed_passes"]
    log(f"timed {args.steps} steps in {elapsed:.3f} s, passes/step={passes}")
    dim = 1 << n
    ms_per_step = elapsed * 1e3 / args.steps
    value = gates * dim * args.steps / elapsed
    # algorithmic state traffic: every pass reads and writes the state, except
    # the first register pass of a run from |0...0>, which only writes it
    step_bytes = (2.0 * passes - (0 if args.lds else 1)) * dim * 8
    bytes_per_pass = step_bytes / max(passes, 1)
    avg_pass_ms = pass_ms / max(timed_passes, 1)
    achieved = bytes_per_pass / (avg_pass_ms * 1e-3) / 1e9
    hbm_frac_step = step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS

    result = {
        "metric": "amp-updates/s (30-qubit complex64 random circuit)",
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"random U3 + CNOT brickwork, {n} qubits, depth {depth}, seed 102, "
                               f"complex64 state, tiled multi-gate passes; each step runs the whole circuit "
                               f"from |0...0> (the first pass generates the start state)",
                   "qubits": n, "depth": depth, "gates": gates, "passes_per_step": passes,
                   "parallelism": "single GPU"},
        "hbm_frac_per_step": hbm_frac_step,
        "passes_compiled": bool(tuned["compile"]) if tuned else bool(run_kw["compile"]),
        "tuned_plan": tuned,
        # first execution of the circuit includes generating and compiling its pass kernels
        "first_run_s": warm_s[0] if warm_s else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args, plan_tile, plan_compiled),
                     "kernel": "k_tile" if args.lds else (
                         "nqpu_pass_* (compiled)" if (tuned["compile"] if tuned else run_kw["compile"])
                         else "k_tile_reg"),
                     "avg_launch_ms": avg_pass_ms,
                     "bytes_per_launch": bytes_per_pass},
    }
    if not args.no_check:
        log("accuracy run: complex128 state of the same circuit")
        ref = E.StateVector(n, "c128")
        t128 = time.perf_counter()
        ref.run(nc, **dict(run_kw, compile=False))  # interpreted: no run-time compilation for a single run
        result["c128_ms_per_step"] = (time.perf_counter() - t128) * 1e3
        # compiled complex128 passes (register/lane exchanges): a run that
        # compiles the pass kernels, then a timed run
        tcc = time.perf_counter()
        ref.reset()
        ref.run(nc, fusion="tiled", compile=True, tile_qubits=11)
        ccompile = time.perf_counter() - tcc
        ref.reset()
        tcc = time.perf_counter()
        stc = ref.run(nc, fusion="tiled", compile=True, tile_qubits=11, time_passes=True)
        result["c128_compiled"] = {"ms_per_step": (time.perf_counter() - tcc) * 1e3, "first_run_s": ccompile,
                                   "passes": stc["passes"], "avg_pass_ms": stc["pass_ms"] / max(stc["timed_passes"], 1),
                                   "exchanges": stc["exchanges"], "remaps": stc["remaps"]}
        sv.reset()
        sv.run(nc, **run_kw)
        ov = sv.overlap(ref)
        result["infidelity_vs_c128"] = 1.0 - abs(ov) ** 2 / (sv.norm() * ref.norm())
        zz_terms = [(0, 3 << i) for i in range(n - 1)]  # all Z_i Z_i+1 in one pass per state
        errs = abs(sv.expectation_pauli_values(zz_terms) - ref.expectation_pauli_values(zz_terms))
        result["zz_max_abs_err_vs_c128"] = float(errs.max())
        result["norm_c64"] = sv.norm()
        ref.close()
    if not args.no_per_gate:
        log("plain per-gate launches on the same GPU")
        sv.reset()
        t1 = time.perf_counter()
        sv.run(nc, fusion="none")
        dt = time.perf_counter() - t1
        result["gpu_per_gate"] = {"amp_updates_per_s": gates * dim / dt, "ms_per_step": dt * 1e3,
                                  "passes": gates}
        result["speedup_vs_per_gate"] = value / result["gpu_per_gate"]["amp_updates_per_s"]
    sv.close()
    log("streaming probe: plain device copy of the same state")
    result["hbm_copy"] = copy_probe(n)
    result["roofline"]["frac_of_copy_rate"] = achieved / result["hbm_copy"]["achieved_gbs"]
    if not args.no_cpu:
        if not args.no_check:
            # the accuracy run uses the headline's plan shape (tuned or fixed)
            acc_kw = run_kw if not tuned else dict(
                fusion="tiled", tile_qubits=tuned["tile_qubits"], low_qubits=tuned["low_qubits"],
                compile=tuned["compile"], fixed_low=tuned["fixed_low"])
            acc, cpu = accuracy_vs_oracle(args, acc_kw)
            result["accuracy_vs_fp64_oracle"] = acc
            result["cpu_baseline"] = cpu
        else:
            log(f"cpu baseline ({args.cpu_seconds} s)")
            result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        result["cpu_baseline"].update(host_info())
        result["speedup_vs_cpu"] = value / result["cpu_baseline"]["value"]
    if not args.only_headline:
        # the other configs at bounded sizes, so the whole line takes a few minutes
        args.batch_steps = args.batch_steps or min(args.steps, 5)
        args.batch_warmup = args.batch_warmup if args.batch_warmup >= 0 else 1
        args.per_circuit_sample = min(args.per_circuit_sample, 256)
        args.qft_steps = args.qft_steps or 2
        args.qft_warmup = args.qft_warmup if args.qft_warmup >= 0 else 1
        args.requests = min(args.requests, 200)
        args.warmup = min(args.warmup, 1)
        args.cpu_seconds = min(args.cpu_seconds, 8.0)
        log("C1: 20-qubit parity circuit")
        result["c1"] = bench_c1(args)
        log("variational loop: plain runs and graph replays")
        result["variational_loop"] = bench_variational_loop(args)
        log("reductions and sampling")
        result["reductions"] = bench_reductions(args)
        log("C3: batched variational circuits")
        result["c3_batch"] = bench_batch(args)
        log("C4: served request stream")
        result["c4_serve"] = bench_serve(args)
        log("C5: memory-ceiling state")
        result["c5_memory_ceiling"] = bench_qft_trotter(args)
        # the sub-results carry their own (redundant) contract keys
        for k in ("c3_batch", "c4_serve", "c5_memory_ceiling"):
            for key in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
                result[k].pop(key, None)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
