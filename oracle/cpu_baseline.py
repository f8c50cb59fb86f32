"""CPU baseline legs of bench.py, run in a child process.

TEST INFRASTRUCTURE ONLY: the reference CPU path is timed here as the
reported baseline (and the oracle computes the FP64 PageRank fixed point the
GPU results are checked against); nothing in the product imports this.

The child process exists so that the OpenMP environment (OMP_NUM_THREADS,
OMP_PLACES=cores, OMP_PROC_BIND=close, as tools/machine/FX700.py:62-64 sets
it for the reference) is in place before libgomp starts; the parent (bench.py)
has torch, and with it an OpenMP runtime, loaded already.  Inputs are the
arrays bench.py generated on the GPU (their digests are the reference's),
passed as raw files in a scratch directory; no GPU is touched here.

    python -m oracle.cpu_baseline spmv DIR       # DIR/{values,rowptr,colidx,x}.bin, DIR/meta.json
    python -m oracle.cpu_baseline pagerank DIR   # DIR/{values,rowptr,colidx}.bin, DIR/meta.json
    python -m oracle.cpu_baseline pagerank-time DIR   # the timing only (no p*)

Each prints one JSON line.
"""
from __future__ import annotations

import hashlib
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def omp_env(threads: int, bind: bool = True) -> dict:
    """The reference's OpenMP setting (bind), or the same thread count with
    the threads left to the OS scheduler (no places, no binding)."""
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    if bind:
        env.update({"OMP_PLACES": "cores", "OMP_PROC_BIND": "close"})
    else:
        env.pop("OMP_PLACES", None)
        env["OMP_PROC_BIND"] = "false"
    return env


def host_cores() -> dict:
    """Threads this run may use (OMP_NUM_THREADS when set, else the CPU
    affinity mask), with nproc and the lscpu model for the record.  Called
    in the parent: the child's OpenMP runtime binds its main thread to one
    core (OMP_PROC_BIND), which narrows its own affinity mask."""
    aff = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", aff))
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = None
    model = platform.processor()
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"threads": threads, "nproc": nproc, "affinity": aff, "cpu_model": model}


def _load(d, meta):
    v = np.fromfile(os.path.join(d, "values.bin"), dtype=np.float64)
    rp = np.fromfile(os.path.join(d, "rowptr.bin"), dtype=np.int32)
    ci = np.fromfile(os.path.join(d, "colidx.bin"), dtype=np.int32)
    assert v.size == meta["n_values"] and rp.size == meta["n_rowptr"] and ci.size == meta["n_colidx"]
    return v, rp, ci


def _stats(ts):
    ts = sorted(ts)
    return {"best_s": ts[0], "median_s": ts[len(ts) // 2], "runs": len(ts)}


def spmv_leg(d: str) -> dict:
    """MatrixCSR::Ax with a preallocated r (the kernel) and spmv() (the API,
    with its n_row-length result allocated and zeroed, MatrixCSR.cpp:123-127),
    1 warm-up and 10 timed repetitions each, on the block bench.py timed."""
    from oracle.oracle import Reference
    meta = json.load(open(os.path.join(d, "meta.json")))
    v, rp, ci = _load(d, meta)
    x = np.fromfile(os.path.join(d, "x.bin"), dtype=np.float64)
    R = Reference
    m = R.from_arrays("CSR", meta["n_row"], meta["n_col"], v, rp, ci)
    del v, rp, ci
    reps = 10
    t, y = R.Ax_times(m, meta["n_row"], x, reps + 1)  # the first call warms up
    ts = R.spmv_times(m, x, reps + 1)
    return {"Ax": _stats(list(t[1:])), "spmv_api": _stats(list(ts[1:])),
            "y_sha256": hashlib.sha256(y.tobytes()).hexdigest()}


def pagerank_leg(d: str, pstar: bool = True) -> dict:
    """The reference's Matrix::page_rank (Matrix.cpp:13-61) timed on a
    bounded number of iterations (its stopping test cannot fire with
    epsilon 0), and the FP64 fixed point p* from the oracle's restatement of
    the same iteration run to its limit cycle, written to DIR/pstar.bin."""
    from oracle.oracle import Oracle, Reference
    meta = json.load(open(os.path.join(d, "meta.json")))
    v, rp, ci = _load(d, meta)
    n = meta["n"]
    out = {}
    k = meta["sample_iterations"]
    if Reference.available():
        r = Reference.from_arrays("CSR", n, n, v, rp, ci)
        r.page_rank(0.0, meta["beta"], 1)  # warm (first touch of the reference's vectors)
        t = time.perf_counter()
        _, its = r.page_rank(0.0, meta["beta"], k)
        dt = time.perf_counter() - t
        del r
        out["reference"] = {"iterations": its, "seconds": dt, "seconds_per_iteration": dt / its}
    if not pstar:
        return out
    o = Oracle.from_arrays("CSR", n, n, v, rp, ci)
    del v, rp, ci
    t = time.perf_counter()
    pstar, its = o.page_rank(0.0, meta["beta"], meta["pstar_iterations"])
    out["pstar"] = {"iterations": its, "seconds": time.perf_counter() - t,
                    "sha256": hashlib.sha256(pstar.tobytes()).hexdigest()}
    pstar.tofile(os.path.join(d, "pstar.bin"))
    return out


def main():
    task, d = sys.argv[1], sys.argv[2]
    res = spmv_leg(d) if task == "spmv" else pagerank_leg(d, pstar=task == "pagerank")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, ROOT)
    main()
