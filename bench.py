#!/usr/bin/env python3
"""Benchmark of the tree-tensor contraction path on one MI355X.

Workload (one "step"): one H|Psi> sweep -- QuTree's Derivative(), i.e.
HamiltonianRepresentation::build (density tree, regularised rho^-1, h-matrices
and hole matrices of every SOP term) followed by LayerDerivative at every node
-- on the 64-mode balanced binary tree: 16 HO-DVR points per leaf, SPF counts
8,10,11,13,14,16 from the bottom layer to the top's children, 64 one-mode terms
T_i + w_i^2 q_i^2 / 2 plus 2016 pair terms c_ij q_i q_j (c_ij ~ N(0, 0.01^2)),
random orthonormal tree state.  Inputs are resident in HBM before timing.

Prints one JSON line.  --gpus above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H|Psi> sweeps/s & CMF steps/s, 64-mode tree, at ≤1e-8 rel. energy error"
PEAK = {"complex128": 78.6, "complex64": 157.3}  # TFLOP/s, dense FP64 / FP32
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_sweep64.md")


def measured_traffic(kernel: str, dtype: str):
    """HBM bytes per launch of `kernel` on the 64-mode sweep, as measured
    with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a copy kernel) and
    recorded in the profile summary; None when it holds no figure."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic_per_launch" and p[1] == kernel and p[2] == dtype:
                    return float(p[3])
    except OSError:
        pass
    return None


def host_cpu():
    """Core count and model of the host CPU (os.cpu_count, /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": avail, "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(text, H, psi, gpu_energy, repeats=5):
    """FP64 CPU restatement of the reference path (the oracle, -O3
    -ffast-math like the reference), timed on this host: one warm-up sweep,
    then `repeats` sweeps, each timed alone with a steady clock; the median.
    Single thread (the reference's default, OpenMP off), then the OpenMP
    build of the same source over the independent nodes and SOP terms of a
    sweep on every core this process may use."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import statistics

    import oracle_lib as O

    def median_rate(variant):
        def run():
            om = O.OracleModel(text, H)
            om.derivative(psi)  # warm-up: page in, build plans
            times = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                om.derivative(psi)
                times.append(time.perf_counter() - t0)
            threads = om._L.orc_threads()
            return 1.0 / statistics.median(times), times, threads, om.energy(psi)
        if variant is None:
            return run()
        with O.build_variant(variant):
            return run()

    one, t1, th1, e_ref = median_rate(None)
    many, tn, thn, _ = median_rate("omp")
    cpu = host_cpu()
    return {"value": one, "unit": "sweeps/s", "cores": th1, "kind": "port",
            "sample": f"{repeats} full sweeps of the same 64-mode workload after one warm-up, median "
                      "(FP64 oracle, -O3 -ffast-math, one thread)",
            "sweep_seconds": [round(x, 4) for x in t1],
            "all_cores": {"value": many, "unit": "sweeps/s", "cores": thn, "sweep_seconds": [round(x, 4) for x in tn],
                          "note": "same oracle built with OpenMP over nodes and SOP terms"},
            "host": cpu}, e_ref


def time_form(Q, text, H, psi, precision, compress, graph, steps, path=1):
    """Sweeps/s and energy of one engine configuration on the same state."""
    eng = Q.Engine(text, H, precision=precision, compress=compress, graph=graph, path=path)
    eng.set_psi(psi)
    eng.derivative(sync=False)
    Q.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.derivative(sync=False)
    Q.synchronize()
    dt = (time.perf_counter() - t0) / steps
    e = eng.energy()
    flops = eng.executed_flops()
    del eng
    return {"sweeps_per_s": round(1.0 / dt, 3), "ms_per_sweep": round(dt * 1e3, 4), "energy": e,
            "executed_flops": flops}


def cmf128_leg(Q, W, steps, chunks=1):
    """The precision run of the 128-mode tree (cfg5's workload): CMF at
    dt = 0.05 with rho regularised at eps = 1e-8 from a random state, in
    complex128 and in the split form (complex64 bulk contractions; FP64 for
    the large terms, the densities, rho^-1, the projector and step control),
    the same number of CMF iterations each: iterations per second and the
    split form's energy and autocorrelation error against complex128 at the
    common output times.  With chunks > 1 each form runs as that many
    integrate calls in a row (each continues from the last state and time;
    the step control restarts at dt), with a progress line per chunk on
    stderr; autocorrelations are then those within each chunk."""
    import numpy as np
    text, H = W.cmf128_model()
    res = {"workload": f"128-mode tree, 128 + 8128 coupled-quartic terms, CMF dt = 0.05, outputs every 0.05, "
                       f"eps 1e-8, {steps} iterations from fill_random(128)"
                       + (f" in {chunks} integrate calls" if chunks > 1 else "")}
    ref = None
    for prec, name in ((0, "complex128"), (2, "complex64_split")):
        eng = Q.Engine(text, H, precision=prec, rho_eps=1e-8, graph=False)
        psi = eng.fill_random(128, delta_lowest=False)
        cmf = Q.CMF(eng)
        cmf.integrate(psi, 0.0, 1000.0, 0.05, 0.05, max_steps=1)  # warm-up
        Q.synchronize()
        r = {"steps": 0, "rejected": 0, "derivs": 0, "times": [], "energies": [], "autocorr": []}
        cur, t, dt = psi, 0.0, 0.0
        for c in range(chunks):
            n = steps // chunks + (steps % chunks if c == chunks - 1 else 0)
            t0 = time.perf_counter()
            rc = cmf.integrate(cur, t, 1000.0, 0.05, 0.05, max_steps=n)
            Q.synchronize()
            dt += time.perf_counter() - t0
            skip = 0 if c == 0 else 1  # a later chunk's first output is the last chunk's end
            for k in ("steps", "rejected", "derivs"):
                r[k] += rc[k]
            for k in ("times", "energies", "autocorr"):
                r[k] += list(rc[k][skip:])
            cur, t = rc["psi"], float(rc["times"][-1])
            if chunks > 1:
                print(f"bench: cmf128 {name} chunk {c + 1}/{chunks} t {t:.3f} steps {r['steps']} "
                      f"{dt:.1f} s", file=sys.stderr, flush=True)
        for k in ("times", "energies", "autocorr"):
            r[k] = np.asarray(r[k])
        out = {"iterations_per_s": round(r["steps"] / dt, 3), "iterations": r["steps"], "rejected": r["rejected"],
               "node_derivatives": r["derivs"], "seconds": round(dt, 3),
               "t_reached": float(r["times"][-1]) if len(r["times"]) else None}
        if len(r["energies"]):
            out["energy_first"] = float(r["energies"][0])
            out["energy_last"] = float(r["energies"][-1])
        if ref is None:
            ref = r
        else:
            n = min(len(ref["times"]), len(r["times"]))
            if n:
                out["energy_rel_err_vs_complex128"] = float(np.max(np.abs(r["energies"][:n] - ref["energies"][:n])
                                                                  / np.abs(ref["energies"][:n])))
                out["autocorr_err_vs_complex128"] = float(np.max(np.abs(r["autocorr"][:n] - ref["autocorr"][:n])))
            out["speedup_vs_complex128"] = round(out["iterations_per_s"] / res["complex128"]["iterations_per_s"], 3)
        res[name] = out
        del cmf, eng
    return res


def gradient_leg(Q, W, gd_steps):
    """Energy gradient dE/dA* on the 32-mode anharmonic model (8 DVR, SPF 8
    then 24): gradient evaluations per second (forward sweep + mean-field
    pass + gradient assembly + copy to the host), then plain gradient
    descent from a random state with the QR gauge after every step, all on
    the device (qtg_engine_gradient_descent)."""
    text, H = W.grad32_model()
    eng = Q.Engine(text, H)
    psi = eng.fill_random(32, delta_lowest=False)
    eng.set_psi(psi)
    eng.gradient()
    n = 10
    t0 = time.perf_counter()
    for _ in range(n):
        eng.gradient()
    t1 = time.perf_counter()
    eng.gradient_descent(psi, 1, 0.05)  # warm-up
    t2 = time.perf_counter()
    _, energies = eng.gradient_descent(psi, gd_steps, 0.05)
    t3 = time.perf_counter()
    energies = [float(x) for x in energies]
    return {"workload": "32-mode quartic model, 8 DVR, SPF 8 then 24 (24x24x24 nodes), 32 + 496 terms",
            "gradient_evals_per_s": round(n / (t1 - t0), 3), "ms_per_gradient": round((t1 - t0) / n * 1e3, 3),
            "gd_steps": gd_steps, "gd_eta": 0.05, "gd_seconds": round(t3 - t2, 3),
            "gd_ms_per_step": round((t3 - t2) / max(gd_steps, 1) * 1e3, 3),
            "gd": "device: forward, backward, step and QR gauge per step, energies copied once",
            "energy_start": energies[0], "energy_end": energies[-1],
            "monotone": bool(all(b < a for a, b in zip(energies, energies[1:])))}


EIG_ARGS = dict(dt=0.05, out=2.0, iterations=3, acc_cmf=1e-7, acc_bs=1e-8)


def eigenstates_leg(Q, W):
    """Imaginary-time eigenstates (Eigenstates.cpp:60-103) of 8 coupled HO
    modes with three states at the top: wall time of the whole run, and the
    levels against the closed form E_0 = sum W_k / 2, E_0 + W_k."""
    import numpy as np

    text, H, K = W.eig8_model(3)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(11, delta_lowest=False)
    cmf = Q.CMF(eng)
    cmf.eigenstates(psi, EIG_ARGS["dt"], EIG_ARGS["out"], 0)  # warm-up
    Q.synchronize()
    t0 = time.perf_counter()
    r = cmf.eigenstates(psi, **EIG_ARGS)
    Q.synchronize()
    dt = time.perf_counter() - t0
    w = np.sqrt(np.linalg.eigvalsh(K))
    exact = np.concatenate([[0.5 * w.sum()], 0.5 * w.sum() + w[:2]])
    return {"workload": "8 coupled HO modes (w_i = 1 + 0.25 i), 4 SPFs, 3 states; "
                        f"{EIG_ARGS['iterations']} iterations of imaginary-time CMF over {EIG_ARGS['out']} a.u., "
                        f"cmf {EIG_ARGS['acc_cmf']}, bs {EIG_ARGS['acc_bs']}",
            "seconds": round(dt, 3), "iterations_per_s": round(EIG_ARGS["iterations"] / dt, 3),
            "cmf_steps": r["steps"], "cmf_steps_per_s": round(r["steps"] / dt, 1),
            "node_derivatives": r["derivs"], "levels": [float(x) for x in r["energies"][-1]],
            "levels_rel_err_vs_closed_form": [float(x) for x in np.abs(r["energies"][-1] - exact) / exact]}, psi


def cpu_eigenstates(psi):
    """The oracle on the same eigenstates run, one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from qutree_amd import workloads as W

    text, H, _ = W.eig8_model(3)
    om = O.OracleModel(text, H)
    t0 = time.perf_counter()
    r = om.eigenstates(psi, **EIG_ARGS)
    dt = time.perf_counter() - t0
    return {"seconds": round(dt, 3), "cmf_steps": r["steps"], "cores": 1, "kind": "port"}


def scf_leg(Q, W, iters=10, cpu=True, eager=False):
    """Krylov-sweep SCF (SCF.cpp:270-348) on the reference's own SCF example
    input, examples/mctdh/CoupledHarmonicOscillator.128.4.4.yaml: 128 coupled
    HO modes on a balanced tree (10 DVR points, 4 SPFs at the bottom rising
    by 4 per layer), Hartree start, nKrylov 10, nIter 10; and the 32-mode
    quartic model (8 DVR, SPF 8 then 24) with nKrylov 5.  Iterations per
    second on the device (graph-replayed), the energy against the closed-form
    ground state of the 128 normal modes, and the oracle timed on one
    iteration of the same run (one thread)."""
    import numpy as np
    from qutree_amd import mctdh as M

    text = M.balanced_tree(128, 10, 4, 4, 0, 0.01822, 0.0, 0.0, 0.01822)
    H = M.coupled_ho(128)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(128, delta_lowest=True)
    scf = Q.SCF(eng, 10)
    if eager:
        scf.set_option("graph", 0)
    scf.run(psi, 1)  # warm-up: plans and captures the iteration graph
    t0 = time.perf_counter()
    r = scf.run(psi, iters)
    dt = time.perf_counter() - t0
    info = scf.info()
    lam, om = 2000.0 / M.CM, 4000.0 / M.CM
    f = 128
    e0 = 0.5 * float(np.sum(np.sqrt(om ** 2 + 2 * lam ** 2 * np.cos(2 * np.pi * np.arange(f) / f))))
    e_it = [float(x) for x in r["energies"][:, -1]]
    out = {"workload": "reference example CoupledHarmonicOscillator.128.4.4.yaml: 128 coupled HO modes, "
                       "10 DVR, SPF 4 + 4 per layer, Hartree start, nKrylov 10",
           "iterations": iters, "seconds": round(dt, 4), "iterations_per_s": round(iters / dt, 3),
           "ms_per_iteration": round(dt / iters * 1e3, 3), "microsteps_per_iteration": info["sweep_length"],
           "launches_per_iteration": info["launches"],
           "us_per_launch": round(dt / iters / info["launches"] * 1e6, 3),
           "energy_per_iteration": e_it, "closed_form_ground_state": e0,
           "rel_gap_to_closed_form": (e_it[-1] - e0) / e0}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        omod = O.OracleModel(text, H)
        c0 = time.perf_counter()
        ref = omod.scf(psi, 1, 10)
        c1 = time.perf_counter()
        out["cpu_oracle"] = {"seconds_per_iteration": round(c1 - c0, 3), "cores": 1, "kind": "port",
                             "sample": "one SCF iteration of the same run"}
        out["speedup_vs_cpu_oracle"] = round((c1 - c0) / (dt / iters), 2)
        out["first_iteration_rel_err_vs_oracle"] = float(
            np.max(np.abs(r["energies"][0] - ref["energies"][0]) / np.abs(ref["energies"][0])))
        # from the Hartree start the unoccupied SPF directions that each QR
        # completes arbitrarily set the next node's search space, so rounding
        # reaches the energies: the same oracle built without fast-math (a
        # second summation order) measures that floor
        with O.strict_build():
            om2 = O.OracleModel(text, H)
            ref2 = om2.scf(psi, 1, 10)
            del om2
        out["oracle_floor_first_iteration"] = float(
            np.max(np.abs(ref2["energies"][0] - ref["energies"][0]) / np.abs(ref["energies"][0])))
    text4, H4 = W.grad32_model()
    eng4 = Q.Engine(text4, H4)
    psi4 = eng4.fill_random(32, delta_lowest=True)
    s4 = Q.SCF(eng4, 5)
    s4.run(psi4, 1)
    t0 = time.perf_counter()
    r4 = s4.run(psi4, 3)
    d4 = time.perf_counter() - t0
    out["quartic32"] = {"workload": "32-mode quartic model, 8 DVR, SPF 8 then 24, nKrylov 5",
                        "ms_per_iteration": round(d4 / 3 * 1e3, 3), "launches_per_iteration": s4.info()["launches"],
                        "energy_per_iteration": [float(x) for x in r4["energies"][:, -1]]}
    return out


def apply_fit_leg(Q, W, cpu=True):
    """The SOP application fitted into a tree (ApplyFirstOrder.cpp:43-147) on
    the 32-mode tree of the gradient workload (8 DVR, SPF 8 then 24): 32
    nearest-neighbour products A_i (x) B_{i+1} of random 8 x 8 leaf matrices
    with spectral norm 0.15, from a random state and a seeded start, with the
    reference's stopping rule (here after 4 iterations, when the fidelity
    falls); iterations per second on the device and the oracle timed on the
    first two iterations of the same run (one thread)."""
    import numpy as np
    text, _ = W.grad32_model()
    rng = np.random.default_rng(5)
    H = W.Hamiltonian()
    for i in range(32):
        mats = []
        for _ in range(2):
            A = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
            mats.append(0.15 * A / np.linalg.norm(A, 2))
        H.add_term(1.0, [(i, W.FACTOR_MATRIX, mats[0]), ((i + 1) % 32, W.FACTOR_MATRIX, mats[1])])
    eng = Q.Engine(text, H)
    psi = eng.fill_random(64, delta_lowest=False)
    fit = Q.ApplyFit(eng)
    fit.run(psi, 3, 2)  # warm-up: plans and captures the iteration graph
    t0 = time.perf_counter()
    r = fit.run(psi, 3, 100)
    dt = time.perf_counter() - t0
    n = len(r["fidelities"])
    out = {"workload": "32-mode tree (8 DVR, SPF 8 then 24), 32 nearest-neighbour products of random 8 x 8 "
                       "matrices (spectral norm 0.15), random state, seed 3",
           "iterations": n, "seconds": round(dt, 4), "iterations_per_s": round(n / dt, 3),
           "ms_per_iteration": round(dt / n * 1e3, 3), "fidelity_final": float(r["fidelities"][-1]),
           "launches_per_iteration": fit.info()["launches"]}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        om = O.OracleModel(text, H)
        c0 = time.perf_counter()
        ref = om.apply_fit(psi, 3, 2)
        c1 = time.perf_counter()
        out["cpu_oracle"] = {"seconds_per_iteration": round((c1 - c0) / len(ref["fidelities"]), 3), "cores": 1,
                             "kind": "port", "sample": "the first two iterations of the same run"}
        out["speedup_vs_cpu_oracle"] = round(out["cpu_oracle"]["seconds_per_iteration"] / (dt / n), 2)
        out["fidelity_rel_err_vs_oracle_first_two"] = float(
            np.max(np.abs(r["fidelities"][:2] - ref["fidelities"]) / np.abs(ref["fidelities"])))
    return out


def big_tree_leg(Q, W, steps=3):
    """One H|Psi> sweep on the 128-mode tree (8 DVR, SPF 8..48 toward the root,
    128 + 8128 coupled-quartic terms) in complex128, complex64 bulk and the
    split form."""
    text, H = W.cmf128_model()
    res = {"workload": "128-mode binary tree, 8 DVR, SPF 8,14,21,27,34,41,48, 128 + 8128 terms"}
    e64 = None
    psi = None
    for prec, name in ((0, "complex128"), (1, "complex64_bulk"), (2, "complex64_split")):
        eng = Q.Engine(text, H, precision=prec)
        if psi is None:
            psi = eng.fill_random(128, delta_lowest=False)
        eng.set_psi(psi)
        eng.derivative(sync=False)
        Q.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.derivative(sync=False)
        Q.synchronize()
        dt = (time.perf_counter() - t0) / steps
        e = eng.energy()
        r = {"sweeps_per_s": round(1.0 / dt, 3), "ms_per_sweep": round(dt * 1e3, 3), "energy": e,
             "executed_flops": eng.executed_flops(), "tflops_per_s_executed": round(eng.executed_flops() / dt / 1e12, 3)}
        if e64 is None:
            e64 = e
            r["reference_flops"] = eng.counts()[0]
        else:
            r["energy_rel_err_vs_complex128"] = abs(e - e64) / abs(e64)
        res[name] = r
        del eng
    return res


def microbench_leg(Q, W, reps=5):
    """Node-level grouped kernels on 4096 order-3 tensors with legs from
    {8,16,24,32,48,64}: matrixTensor on a random mode and the hole
    contraction(A, B, k), complex128, with their rooflines."""
    import numpy as np

    shapes, modes = W.microbench_shapes()
    count = len(modes)
    sizes = np.prod(shapes, axis=1)
    offT = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    nk = shapes[np.arange(count), modes]
    offM = np.concatenate([[0], np.cumsum(nk * nk)[:-1]]).astype(np.int64)
    tot, totM = int(sizes.sum()), int((nk * nk).sum())
    mp = np.zeros(count, Q.MP_DESC)
    hc = np.zeros(count, Q.HC_DESC)
    for t in range(count):
        d, k = shapes[t], modes[t]
        before, after = int(np.prod(d[:k])), int(np.prod(d[k + 1:]))
        mp[t] = (offT[t], offM[t], offT[t], before, d[k], d[k], after, 0, 0)
        hc[t] = (offM[t], offT[t], offT[t], before, d[k], d[k], after, 0, 0)
    # synthetic values (timing only; parity is covered by the test suite)
    rng = np.random.default_rng(4096)
    block = rng.standard_normal(2 * 4096)
    fill = np.resize(block, 2 * tot).view(np.complex128)
    dA = Q.DeviceBuffer.from_array(fill)
    dB = Q.DeviceBuffer.from_array(fill)
    dC = Q.DeviceBuffer(16 * tot)
    dM = Q.DeviceBuffer.from_array(np.resize(block, 2 * totM).view(np.complex128))
    dS = Q.DeviceBuffer(16 * totM)
    out = {"workload": f"{count} order-3 tensors, legs from {{8,16,24,32,48,64}}, complex128",
           "elements": tot}
    for name, plan, args in (("matrixTensor", Q.Plan("mode_product", mp), (dC, dM, dA)),
                             ("contraction", Q.Plan("hole", hc), (dS, dA, dB))):
        plan.execute(*args)
        Q.synchronize()
        tm = Q.Timer()
        tm.start()
        for _ in range(reps):
            plan.execute(*args)
        ms = tm.stop() / reps
        fl, by = plan.counts()
        tf, gbs = fl / (ms * 1e-3) / 1e12, by / (ms * 1e-3) / 1e9
        t_roof = max(by / 8e12, fl / 78.6e12)
        out[name] = {"ms": round(ms, 4), "tflops_per_s": round(tf, 3), "gb_per_s": round(gbs, 1),
                     "flops": fl, "bytes": by, "roofline_frac": round(t_roof / (ms * 1e-3), 4),
                     "bound": "hbm" if by / 8e12 > fl / 78.6e12 else "mfma"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["complex128", "complex64"], default="complex128")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cmf-steps", type=int, default=60,
                    help="CMF iterations (Integrate loop passes, accepted or rejected) timed on the 64-mode tree "
                         "(0: skip)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip timing the plain per-term complex128 form and the complex64 bulk form")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the gradient, 128-mode and node-microbench legs")
    ap.add_argument("--gd-steps", type=int, default=200, help="gradient-descent steps of the gradient leg")
    ap.add_argument("--cmf128-steps", type=int, default=20,
                    help="CMF iterations of the 128-mode precision leg (0: skip; the full run is 1000)")
    ap.add_argument("--cmf128-chunks", type=int, default=1,
                    help="run the 128-mode precision leg as this many integrate calls (progress on stderr)")
    ap.add_argument("--cmf128-only", action="store_true",
                    help="only run the 128-mode precision leg and print its JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the HBM byte counters)")
    ap.add_argument("--scf-only", action="store_true",
                    help="only run the SCF leg on the device with eager launches (no CPU oracle), for profiling")
    args = ap.parse_args()
    if args.scf_only:
        # eager launches: the profiler crashed while tracing the captured
        # iteration graph (26k launches); kernel durations are the same
        import qutree_amd as Q
        from qutree_amd import workloads as W
        print(json.dumps({"scf": scf_leg(Q, W, iters=3, cpu=False, eager=True)}))
        return 0
    if args.cmf128_only:
        import qutree_amd as Q
        from qutree_amd import workloads as W
        print(json.dumps({"cmf128": cmf128_leg(Q, W, args.cmf128_steps, args.cmf128_chunks)}))
        return 0
    if args.copy_probe > 0:
        import ctypes as C
        import qutree_amd as Q
        L = Q.lib()
        n = args.copy_probe << 20
        src, dst = C.c_void_p(), C.c_void_p()
        Q.check(L.qtg_device_alloc(C.byref(src), n))
        Q.check(L.qtg_device_alloc(C.byref(dst), n))
        L.qtg_copy_d.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        for _ in range(5):  # the library's copy kernel: reads n bytes, writes n bytes
            Q.check(L.qtg_copy_d(dst, src, n // 8, None))
        Q.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "copies": 5}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding over SOP terms is design-only", file=sys.stderr)
        return 2

    import numpy as np

    import qutree_amd as Q
    from qutree_amd import workloads as W

    text, H = W.sweep64_model()
    prec = 0 if args.precision == "complex128" else 1
    eng = Q.Engine(text, H, precision=prec, graph=True)
    psi = eng.fill_random(64, delta_lowest=False)
    eng.set_psi(psi)
    for _ in range(args.warmup):
        eng.derivative(sync=False)
    Q.synchronize()

    # timed region: K sweeps replayed from the captured HIP graph
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.derivative(sync=False)
    Q.synchronize()
    t1 = time.perf_counter()
    sweep_s = (t1 - t0) / args.steps
    value = 1.0 / sweep_s

    # the same sweep with host buffers at the boundary: the state copied in
    # from host memory and the derivative copied back out, per sweep (the
    # PCIe-inclusive rate; not the headline value)
    n_host = max(5, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(n_host):
        eng.set_psi(psi)
        eng.derivative(sync=False)
        eng.get(0)
    t1 = time.perf_counter()
    host_rate = n_host / (t1 - t0)

    # the same K sweeps launched eagerly with HIP events around every launch
    # on the engine's streams (each sweep queued behind a short device-side
    # hold, so the events time kernels rather than host launch gaps):
    # per-kernel durations for the roofline
    eng.set_timing(True)
    Q.synchronize()
    for _ in range(args.steps):
        eng.derivative(sync=False)
    Q.synchronize()
    timing = eng.read_timing()
    eng.set_timing(False)
    energy = eng.energy()
    flops, items = eng.counts()  # reference-algorithm count of one sweep
    executed = eng.executed_flops()

    # dominant kernel = largest summed event time among the kernels of the
    # engine stream (the eigensolver runs on the side stream, overlapped with
    # the represent and hole launches, and is reported under "kernels"); its
    # bound is whichever of the algorithmic FLOPs (at the MFMA peak) and
    # bytes (at the HBM peak) takes longer
    side_stream = {"heev_reginv_kernel"}
    dom = max((k for k in timing if k not in side_stream), key=lambda k: timing[k]["ms"])
    d = timing[dom]
    nl = max(1, d["launches"])
    avg_ms = d["ms"] / nl
    flops_per_launch = d["flops"] / nl
    bytes_per_launch = d["bytes"] / nl
    peak_tf = PEAK[args.precision]
    if bytes_per_launch / (HBM_PEAK_GBS * 1e9) > flops_per_launch / (peak_tf * 1e12):
        bound, achieved, peak, unit = "hbm", bytes_per_launch / (avg_ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        bound, achieved, peak, unit = "mfma", flops_per_launch / (avg_ms * 1e-3) / 1e12, peak_tf, "TFLOP/s"
    roofline = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 5), "traffic": measured_traffic(dom, args.precision), "kernel": dom,
                "avg_launch_ms": round(avg_ms, 5), "launches_per_step": d["launches"] / args.steps,
                "flops_per_launch": flops_per_launch, "algorithmic_bytes_per_launch": bytes_per_launch}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sweep_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": "H|Psi> sweep (Derivative: build + LayerDerivative at all nodes), 64-mode "
                               "balanced binary tree, 16 HO-DVR, SPF 8,10,11,13,14,16, 64 one-mode + 2016 "
                               "pair terms",
                   "modes": 64, "nodes": eng.nnodes, "terms": H.nterms, "psi_elements": int(eng.psi_size),
                   "parallelism": "single GPU"},
        "sweep_flops_reference": flops,
        "sweep_flops_executed": executed,
        "tflops_per_s_reference_equivalent": round(flops * value / 1e12, 3),
        "tflops_per_s_executed": round(executed * value / 1e12, 3),
        "kernels": {k: {"ms_per_step": round(v["ms"] / args.steps, 4),
                        "launches_per_step": v["launches"] / args.steps} for k, v in timing.items()},
        "energy": energy,
        "host_buffers": {"sweeps_per_s": round(host_rate, 3),
                         "per_sweep": "state in from host memory, derivative out to host memory "
                                      f"({16 * int(eng.psi_size)} bytes each way)"},
        "roofline": roofline,
    }
    if not args.no_forms:
        # the plain drop-in form (one work item per (term, node) as the
        # reference loops, complex128, eager launches) and the complex64 bulk
        # form, on the same state; energies relative to this run's complex128
        plain = time_form(Q, text, H, psi, 0, False, False, max(3, args.steps // 10), path=0)
        lds = time_form(Q, text, H, psi, 0, True, True, args.steps, path=0)
        c64 = time_form(Q, text, H, psi, 1, True, True, args.steps)
        split = time_form(Q, text, H, psi, 2, True, True, args.steps)
        for f in (plain, lds, c64, split):
            f["energy_rel_err_vs_complex128"] = abs(f["energy"] - energy) / abs(energy)
        plain["form"] = "complex128, per-term work items, whole node tensors in LDS, eager launches"
        lds["form"] = "complex128, shared slots, whole node tensors in LDS, HIP graph"
        c64["form"] = ("complex64 bulk MFMA contractions (FP64 rho^-1, projector, output), shared slots, "
                       "HIP graph")
        split["form"] = ("complex64 bulk MFMA contractions, FP64 for the large-coefficient terms and the "
                         "densities (and rho^-1, projector, output), shared slots, HIP graph")
        out["forms"] = {"plain_complex128": plain, "lds_complex128": lds, "complex64_bulk": c64,
                        "complex64_split": split,
                        "speedup_vs_plain": round(value / plain["sweeps_per_s"], 3),
                        "complex64_speedup_vs_plain": round(c64["sweeps_per_s"] / plain["sweeps_per_s"], 3),
                        "complex64_split_speedup_vs_plain": round(split["sweeps_per_s"] / plain["sweeps_per_s"], 3)}
    if args.cmf_steps > 0 and args.precision == "complex128":
        # CMF steps (CMFIntegrator::Integrate iterations: 5 half/quarter
        # substeps of per-node Bulirsch-Stoer, 2-3 builds, error control,
        # orthogonal) on the same tree from a displaced Hartree product
        text_c, H_c = W.cmf64_model()

        def cmf_form(precision, path, compress):
            engc = Q.Engine(text_c, H_c, precision=precision, compress=compress, graph=False, path=path)
            psi_c = engc.fill_random(64, delta_lowest=True)
            cmf = Q.CMF(engc)
            cmf.integrate(psi_c, 0.0, 10.0, 0.1, 0.1, max_steps=1)  # warm-up
            Q.synchronize()
            c0 = time.perf_counter()
            r = cmf.integrate(psi_c, 0.0, 10.0, 0.1, 0.1, max_steps=args.cmf_steps)
            Q.synchronize()
            c1 = time.perf_counter()
            acc = r["steps"] - r["rejected"]
            return {"steps_per_s": round(r["steps"] / (c1 - c0), 3),
                    "iterations_per_s": round(r["steps"] / (c1 - c0), 3),
                    "accepted_steps_per_s": round(acc / (c1 - c0), 3), "iterations": r["steps"],
                    "steps": r["steps"], "accepted": acc,
                    "rejected": r["rejected"], "builds": r["builds"], "node_derivatives": r["derivs"],
                    "ticks": r["ticks"], "ms_per_iteration": round((c1 - c0) / r["steps"] * 1e3, 3),
                    "seconds": round(c1 - c0, 4), "t_reached": float(r["times"][-1]) if len(r["times"]) else None,
                    "energy_final": float(r["energies"][-1]) if len(r["energies"]) else None}, r

        main_cmf, r64 = cmf_form(0, 1, True)
        main_cmf["workload"] = "64-mode tree, displaced Hartree start (q0 = 1), dt = 0.1, cmf 1e-4, bs 1e-5"
        out["cmf"] = main_cmf
        if not args.no_forms:
            plain_cmf, _ = cmf_form(0, 0, False)
            split_cmf, rs = cmf_form(2, 1, True)
            e_ref = np.asarray(r64["energies"])
            n = min(len(e_ref), len(rs["energies"]))
            split_cmf["energy_rel_err_vs_complex128"] = (
                float(np.max(np.abs(np.asarray(rs["energies"][:n]) - e_ref[:n]) / np.abs(e_ref[:n]))) if n else None)
            out["cmf_forms"] = {"plain_complex128": plain_cmf, "complex64_split": split_cmf,
                                "speedup_vs_plain": round(main_cmf["steps_per_s"] / plain_cmf["steps_per_s"], 3),
                                "complex64_split_speedup_vs_plain":
                                    round(split_cmf["steps_per_s"] / plain_cmf["steps_per_s"], 3)}
    if not args.no_extras and args.precision == "complex128":
        def leg(name):  # a progress line per leg on stderr (stdout keeps the one JSON line)
            print(f"bench: {name} done", file=sys.stderr, flush=True)

        out["gradient"] = gradient_leg(Q, W, args.gd_steps)
        leg("gradient")
        out["tree128"] = big_tree_leg(Q, W)
        leg("tree128")
        if args.cmf128_steps > 0:
            out["cmf128"] = cmf128_leg(Q, W, args.cmf128_steps, args.cmf128_chunks)
            leg("cmf128")
        out["node_microbench"] = microbench_leg(Q, W)
        out["eigenstates"], eig_psi = eigenstates_leg(Q, W)
        leg("microbench, eigenstates")
        out["scf"] = scf_leg(Q, W, cpu=not args.no_cpu)
        leg("scf")
        out["apply_fit"] = apply_fit_leg(Q, W, cpu=not args.no_cpu)
        leg("apply_fit")
    if not args.no_cpu:
        base, e_ref = cpu_baseline(text, H, psi, energy)
        out["cpu_baseline"] = base
        out["energy_rel_err_vs_fp64_oracle"] = abs(energy - e_ref) / abs(e_ref)
        if "eigenstates" in out:
            ce = cpu_eigenstates(eig_psi)
            out["eigenstates"]["cpu_oracle"] = ce
            out["eigenstates"]["speedup_vs_cpu_oracle"] = round(ce["seconds"] / out["eigenstates"]["seconds"], 2)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
