"""Benchmark: converged SCFT solves per second on one MI355X.

Workload (one "step" = one converged solve): AB diblock double gyroid
(Ia-3d), N=100, blocks 36/64, chi N=20, b=1, ds=0.5 (200 contour steps),
cubic cell starting at a=16, 64^3 grid, seeded initial field (first 20 stars
~ N(0, 0.1) per monomer, seed 20261201), Anderson mixing (N_hist=20) with
unit-cell relaxation through scf_stress to residual 1e-9 -- the reference's
ITERATE/iterate_AM path on the GPU hot path.

Prints one JSON line with the headline value, the roofline of the fused
contour step (HIP-event timing of its kernels inside the timed region) and
the reference Fortran+MKL-FFTW program timed on the host (cpu_baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "gyroid_Ia-3d_64^3_N100_f0.36_chiN20_ds0.5_a16_AM_domain"


def load_workload():
    d = np.load(os.path.join(ROOT, "tests", "golden", "gyr64.npz"))
    return {k: d[k] for k in d.files}


def gyroid_system():
    from pscf_amd.system import diblock
    return diblock(0.36, 100.0, 20.0, dim=3, crystal_system="cubic", cell_param=[16.0],
                   ngrid=[64, 64, 64], chain_step=0.5, group_name="I a -3 d")


def step_bytes(M, Mk):
    """Algorithmic HBM bytes of one contour step of one sweep (streaming-FFT
    model: every 3D transform reads its input and writes its output once,
    pointwise factors fused into the adjacent transform)."""
    return 96.0 * M + 120.0 * Mk


def cpu_baseline(g, n_gpu_iter, workdir):
    """The reference program (oracle/_ref/pscf, flang -O2 + MKL's FFTW3 API,
    one core) on the same workload, bounded to its first Anderson iterations.
    Reports its per-iteration cost scaled by the GPU solve's iteration count."""
    ref = os.path.join(ROOT, "oracle", "_ref", "pscf")
    if not os.path.exists(ref):
        return None
    from pscf_amd import fieldio, scf
    sysm = gyroid_system()
    B = scf.Basis.from_tables(g)
    fieldio.write_field(os.path.join(workdir, "in.omega"), g["omega"], sysm, B, g["star_count"])
    max_itr = 1
    param = sysm.to_param() + (
        "ITERATE\ninput_filename\n'in.omega'\noutput_prefix\n'out/'\nmax_itr\n%d\n"
        "error_max\n1.0E-09\ndomain\nT\nitr_algo\n'AM'\nN_hist\n20\n\nFINISH\n" % max_itr)
    with open(os.path.join(workdir, "param"), "w") as f:
        f.write(param)
    os.makedirs(os.path.join(workdir, "out"), exist_ok=True)
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    t0 = time.perf_counter()
    with open(os.path.join(workdir, "param")) as fin:
        r = subprocess.run(["bash", "-c", "ulimit -s unlimited; %s" % ref], stdin=fin,
                           cwd=workdir, env=env, capture_output=True, text=True, timeout=600)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        return {"error": "reference run failed: " + r.stderr[-300:]}
    out = open(os.path.join(workdir, "out", "out")).read()
    scf_time = float(re.search(r"^scf_time\s*\n\s*(\S+)", out, re.M).group(1))
    itr = int(re.search(r"^iterations\s*\n\s*(\S+)", out, re.M).group(1))
    # scf_time covers the initial density() plus `itr` iterations of
    # density() + scf_stress(): per-iteration cost ~ scf_time / (itr + 1)
    t_iter = scf_time / (itr + 1)
    return {"value": 1.0 / (t_iter * n_gpu_iter), "unit": "solves/s", "cores": 1,
            "kind": "reference",
            "sample": "reference pscf (flang -O2, MKL FFTW3 API, MKL_NUM_THREADS=1), same "
                      "workload, ITERATE AM max_itr=%d: scf_time %.2f s for %d iterations -> "
                      "%.3f s/iteration x %d GPU iterations per solve (wall %.1f s incl. "
                      "make_basis)" % (max_itr, scf_time, itr, t_iter, n_gpu_iter, wall)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent solves shard across GPUs without communication; see DESIGN.md)")
        sys.exit(2)

    import torch
    from pscf_amd import iterate, scf
    from pscf_amd._lib import lib

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    L = lib()
    g = load_workload()
    sysm = gyroid_system()
    B = scf.Basis.from_tables(g)
    _, G0, _ = iterate.unit_cell(L, 3, "cubic", [16.0])
    S = scf.SCFSystem(sysm, B, G0, lib=L)
    am = iterate.AndersonMixing(S, n_hist=20, domain=True)

    def solve():
        S.set_cell(G0)
        return am.run(g["omega"], error_max=1e-9, max_itr=2000)

    for _ in range(args.warmup):
        r = solve()
    torch.cuda.synchronize()
    S.profile_sweeps(1)            # HIP events around every captured sweep
    t0 = time.perf_counter()
    results = [solve() for _ in range(args.steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms, n = S.profile_sweeps(0)
    r = results[-1]
    iters = [x.itr for x in results]

    M = 64 ** 3
    Mk = 33 * 64 * 64
    # slot 4: milliseconds of whole sweeps and the contour steps they advanced
    steps_timed = int(n[4])
    t_step_ms = float(ms[4]) / max(steps_timed, 1)
    bytes_step = 2 * step_bytes(M, Mk)           # qf and qr sweeps advance together
    achieved = bytes_step / (t_step_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        wd = tempfile.mkdtemp(prefix="pscf_cpu_")
        try:
            cpu = cpu_baseline(g, int(np.mean(iters)), wd)
        finally:
            shutil.rmtree(wd, ignore_errors=True)

    line = {
        "metric": "converged SCFT solves/s at 64^3 FP64",
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "grid": [64, 64, 64], "contour_steps": 200,
                   "iterations_per_solve": iters, "converged": [x.converge for x in results],
                   "final_cell_a": float(r.cell_param[0]), "f_helmholtz": r.f_helmholtz,
                   "ms_per_iteration": elapsed / sum(iters) * 1e3},
        "roofline": {"bound": "hbm", "kernel": "contour step (4 fused passes, both sweeps)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "us_per_step": t_step_ms * 1e3, "steps_timed": steps_timed,
                     "bytes_per_step": bytes_step},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
