"""Autotuner of the contour-step kernel plans (pscf_amd/csrc/tuned.hpp).

The plan of a grid length n is (max_radix, xpad, zsplit, alias, ajobs): the
largest Stockham radix of its transforms, the LDS padding of the XY-plane
rows, the number of workgroups per ky row of the z-column step kernel,
whether the packed-row and half-spectrum buffers of an XY plane share LDS,
and the workgroups per plane of k_step_a (0: chosen at launch from the
number of planes, CUs and resident workgroups).  The candidate space below keeps
the dimensions that mattered in earlier sweeps (row padding always won,
shared buffers were neutral; both stay in the table format).  They are compile-time
parameters, so the tuner builds one library per candidate plan (overriding
the table for every length), times the captured contour sweeps of a
synthetic system on each candidate and grid length on the device, and writes
the fastest plan per length into tuned.hpp.

    python -m pscf_amd.autotune build              # compile the candidates (CPU)
    python -m pscf_amd.autotune measure [--sizes 32,48,64,96]   # time them (GPU)
    python -m pscf_amd.autotune write              # regenerate csrc/tuned.hpp

Timing workload per length n: an AB diblock (N=100, f=0.3, chi N=20, ds=1,
100 contour steps) in a cubic cell on an n^3 grid with a seeded random omega,
both sweeps per launch, graph-captured, HIP events around every sweep.
"""
import argparse
import ctypes
import itertools
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE_DIR = os.path.join(HERE, "_tune")
RESULTS = os.path.join(TUNE_DIR, "results.json")
TUNED_HPP = os.path.join(HERE, "csrc", "tuned.hpp")

RADIX = (8, 4)
XPAD = (1,)
ZSPLIT = (1, 3)
ALIAS = (0,)
AJOBS = (0, 2, 1)


def candidates():
    return [dict(max_radix=r, xpad=p, zsplit=z, alias=a, ajobs=aj)
            for r, p, z, a, aj in itertools.product(RADIX, XPAD, ZSPLIT, ALIAS, AJOBS)]


def variant_name(c):
    return "r%d_p%d_z%d_a%d_j%d" % (c["max_radix"], c["xpad"], c["zsplit"], c["alias"], c["ajobs"])


def build(jobs=4):
    from .build import build_lib
    os.makedirs(TUNE_DIR, exist_ok=True)

    def one(c):
        out = os.path.join(TUNE_DIR, "lib_%s.so" % variant_name(c))
        build_lib(defines={"PSCF_MAX_RADIX": c["max_radix"], "PSCF_XPAD": c["xpad"],
                           "PSCF_ZSPLIT": c["zsplit"], "PSCF_ALIAS": c["alias"],
                           "PSCF_AJOBS": c["ajobs"]}, out=out)
        return out

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(one, candidates()):
            print("built", out, flush=True)


def _trivial_basis():
    from .scf import Basis
    return Basis(np.zeros((1, 3)), np.ones(1), np.ones(1), np.zeros(1), np.zeros((1, 3)))


def time_sweeps(lib, n, reps=5, seed=7):
    """Microseconds per contour step (both sweeps) of the synthetic n^3 system."""
    import torch
    from .iterate import unit_cell
    from .scf import SCFSystem
    from .system import diblock
    s = diblock(0.3, 100.0, 20.0, dim=3, crystal_system="cubic", cell_param=[float(n) / 3.0],
                ngrid=[n, n, n], chain_step=1.0, group_name="P 1")
    _, G, _ = unit_cell(lib, 3, "cubic", s.cell_param)
    S = SCFSystem(s, _trivial_basis(), G, lib=lib)
    rng = np.random.default_rng(seed)
    dev = lib.tensor_device
    om = torch.tensor(0.2 + 0.1 * rng.standard_normal((2, n, n, n)), device=dev)
    rho = torch.empty_like(om)
    q = np.zeros(1)
    qs = np.zeros(1)
    from ._lib import dptr

    def dens():
        lib.call("pscf_density_grid_dev", S.h, ctypes.c_void_p(om.data_ptr()),
                 ctypes.c_void_p(rho.data_ptr()), dptr(q), dptr(qs), dptr(S.phi_chain),
                 dptr(S.mu_chain), dptr(S.phi_solvent), dptr(S.mu_solvent))

    if dev == "cuda":
        torch.cuda.synchronize()
    dens()
    S.profile_sweeps(1)
    for _ in range(reps):
        dens()
    ms, cnt = S.profile_sweeps(0)
    return float(ms[4]) / max(int(cnt[4]), 1) * 1e3, float(rho.sum().item())


def measure(sizes, reps=5, results=RESULTS):
    from ._lib import Lib
    res = {}
    if os.path.exists(results):
        res = json.load(open(results))
    for c in candidates():
        path = os.path.join(TUNE_DIR, "lib_%s.so" % variant_name(c))
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run `python -m pscf_amd.autotune build` first)")
        lib = Lib(path)
        for n in sizes:
            t0 = time.perf_counter()
            us, chk = time_sweeps(lib, n, reps)
            res.setdefault(str(n), {})[variant_name(c)] = us
            print("%-10s n=%-3d %8.2f us/step  (checksum %.12e, %.1fs)" %
                  (variant_name(c), n, us, chk, time.perf_counter() - t0), flush=True)
        json.dump(res, open(results, "w"), indent=1, sort_keys=True)
    return res


def write(results=RESULTS):
    res = json.load(open(results))
    rows = []
    for n in sorted(res, key=int):
        best = min(res[n], key=res[n].get)
        r, p, z, a, aj = [int(t[1:]) for t in best.split("_")]
        rows.append((int(n), r, p, z, a, aj, res[n][best], res[n].get("r8_p1_z1_a0_j2")))
    lines = [
        "// Kernel plans per FFT/grid length for gfx950, written by pscf_amd/autotune.py",
        "// from timings of the contour sweeps on an MI355X (rerun it to regenerate).",
        "//   max_radix  largest Stockham radix of a transform of this length (8 or 4)",
        "//   xpad       LDS padding (complex words) per packed row of an XY plane",
        "//   zsplit     workgroups sharing one ky row of the z-column step kernel",
        "//   alias      1: the packed-row and half-spectrum LDS buffers of an XY plane",
        "//              share memory (half the LDS per workgroup, more workgroups per CU)",
        "//   ajobs      workgroups per plane of k_step_a: 2 (each repeats the inverse",
        "//              transform and starts one sub-step), 1 (starts both), or 0: 2 while",
        "//              one per plane would leave CUs idle and two fit in one round of",
        "//              resident workgroups, else 1",
        "#pragma once", "", "namespace pscf {", "",
        "struct TunedPlan {", "  int n, max_radix, xpad, zsplit, alias, ajobs;", "};", "",
        "constexpr TunedPlan kTunedPlans[] = {"]
    for n, r, p, z, a, aj, us, base in rows:
        note = "  // %.2f us/step (radix 8, pad 1, 1 row, separate, 2 jobs: %s)" % (
            us, "%.2f" % base if base is not None else "n/a")
        lines.append("    {%d, %d, %d, %d, %d, %d},%s" % (n, r, p, z, a, aj, note))
    lines += ["};", "", "}  // namespace pscf", ""]
    open(TUNED_HPP, "w").write("\n".join(lines))
    print("wrote", TUNED_HPP)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pscf_amd.autotune")
    ap.add_argument("cmd", choices=["build", "measure", "write"])
    ap.add_argument("--sizes", default="32,48,64,96")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--results", default=RESULTS, help="timing table (json)")
    a = ap.parse_args(argv)
    if a.cmd == "build":
        build(a.jobs)
    elif a.cmd == "measure":
        measure([int(v) for v in a.sizes.split(",")], a.reps, a.results)
    else:
        write(a.results)


if __name__ == "__main__":
    sys.exit(main())
