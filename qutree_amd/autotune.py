"""Autotuner of the grouped sweep plan (run on the GPU).

For each workload class it searches the plan parameters that shape the
grouped launches:
  split_waves       hole-contraction launches with fewer tiles split their
                    reductions along K toward this many wavefronts;
  split_min_k       the shortest K chunk of such a split;
  gc_long_k         hole contractions longer than 2 gc_long_k always split;
  gm_split_k        mode products longer than 2 gm_split_k split along K;
by coordinate descent (two passes over the parameters, every candidate of
one parameter with the others fixed), each candidate timed as the median of
three runs of `--sweeps` graph-replayed sweeps on the same state.  The
winning values are written as one line per class into the tuning table
(qutree_amd/tuning/gfx950.txt by default), which the engine reads when a
Hamiltonian is planned (csrc/tuning.cpp):
  class <precision> <max SPF> <log2 state elements> <log2 terms> key value ...
Every candidate must reproduce the untuned energy (relative 1e-12 in
complex128, 1e-7 in the complex64 forms): the parameters change the
summation order only.

Usage: python -m qutree_amd.autotune [--workloads sweep64,tree128,...] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "gfx950.txt")

GRID = {
    "split_waves": [512, 1024, 2048, 4096, 8192],
    "split_min_k": [8, 16, 32, 64],
    "gc_long_k": [128, 256, 512, 1024],
    "gm_split_k": [16, 32, 64],
}
DEFAULTS = {"split_waves": 2048, "split_min_k": 16, "gc_long_k": 512, "gm_split_k": 32}


def workloads():
    from . import workloads as W
    return {
        "sweep64": (W.sweep64_model, 0, 64),
        "sweep64_split": (W.sweep64_model, 2, 64),
        "sweep64_c64": (W.sweep64_model, 1, 64),
        "grad32": (W.grad32_model, 0, 32),
        "tree128": (W.cmf128_model, 0, 128),
        "tree128_split": (W.cmf128_model, 2, 128),
    }


def time_sweeps(text, H, psi, precision, params, sweeps, repeats=3):
    import qutree_amd as Q
    eng = Q.Engine(text, H, precision=precision, tuning=0, **params)
    eng.set_psi(psi)
    for _ in range(3):
        eng.derivative(sync=False)
    Q.synchronize()
    runs = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(sweeps):
            eng.derivative(sync=False)
        Q.synchronize()
        runs.append((time.perf_counter() - t0) / sweeps)
    e = eng.energy()
    del eng
    return statistics.median(runs) * 1e6, e


def class_key(text, H, precision):
    import qutree_amd as Q
    eng = Q.Engine(text, H, precision=precision, tuning=0)
    nspf = max(n["dims"][-1] for n in eng.nodes)
    key = (precision, nspf, int(math.log2(eng.psi_size)), int(math.log2(max(H.nterms, 1))))
    del eng
    return key


def tune(name, sweeps, log=print):
    import qutree_amd as Q
    model, precision, seed = workloads()[name]
    text, H = model()[:2]
    eng = Q.Engine(text, H)
    psi = eng.fill_random(seed, delta_lowest=False)
    del eng
    tol = 1e-12 if precision == 0 else 1e-7
    base_us, e0 = time_sweeps(text, H, psi, precision, dict(DEFAULTS), sweeps)
    best = dict(DEFAULTS)
    best_us = base_us
    log(json.dumps({"workload": name, "untuned_us": round(base_us, 2)}))
    for _ in range(2):
        for key, values in GRID.items():
            for v in values:
                if v == best[key]:
                    continue
                trial = dict(best, **{key: v})
                us, e = time_sweeps(text, H, psi, precision, trial, sweeps)
                if abs(e - e0) > tol * abs(e0):
                    raise RuntimeError(f"{name}: {trial} changed the energy by {abs(e - e0) / abs(e0):.2e}")
                log(json.dumps({"workload": name, key: v, "us": round(us, 2)}))
                if us < best_us * 0.99:
                    best, best_us = trial, us
    return class_key(text, H, precision), best, base_us, best_us


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="sweep64,sweep64_split,grad32,tree128")
    ap.add_argument("--sweeps", type=int, default=30)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    args = ap.parse_args(argv)
    lines = ["# Tuning table of the grouped sweep plan on gfx950 (MI355X), written by",
             "# python -m qutree_amd.autotune; read by csrc/tuning.cpp when a Hamiltonian is planned.",
             "# class <precision> <max SPF> <log2 state elements> <log2 terms> key value ...  # workload, "
             "untuned and tuned microseconds per sweep"]
    for name in args.workloads.split(","):
        key, best, base_us, best_us = tune(name, args.sweeps if not name.startswith("tree128") else 5)
        kv = " ".join(f"{k} {v}" for k, v in best.items())
        lines.append(f"class {key[0]} {key[1]} {key[2]} {key[3]} {kv}  # {name}: untuned {base_us:.1f} us, "
                     f"tuned {best_us:.1f} us")
        print(lines[-1], flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
