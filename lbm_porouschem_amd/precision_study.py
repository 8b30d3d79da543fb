"""Per-field error of reduced-precision storage against an FP64 run of the same seed.

    python -m lbm_porouschem_amd.precision_study [--shape 512,256,256] [--steps 2000]
                                                 [--precisions fp16,fp32] [--collision mrt]

Runs the 3D coupled pyrolysis workload (workloads.coupled3d) on the GPU in FP64 and
in each requested storage precision, then prints one JSON line per precision with
the relative L2 error of every macroscopic field against the FP64 run, the FP64
global balances (fluid mass, fluid and solid heat content, FIX-species mass = the
solid-conversion integral) of both runs, and the MLUPS of each run.  Populations
of the fp16 run are stored as FP16 deviations from the lattice reference values;
all arithmetic and every reduction accumulate in FP32 / FP64 respectively.
"""
import argparse
import json
import sys
import time

import numpy as np


def _fields(lb):
    out = {"rho": lb.rho.to_numpy(), "v": lb.v.to_numpy(), "TF": lb.TF.S.to_numpy(), "TS": lb.TS.S.to_numpy(),
           "rhos": lb.rhos.to_numpy()}
    for s in lb.species:
        out["S_" + s.name] = s.S.to_numpy()
    return out


def _balances(lb):
    from . import _abi
    lib, h = lb._lib, lb._handle
    lb._flush()
    return {"mass": lib.reduce(h, _abi.REDUCE_MASS), "TF_sum": lib.reduce(h, _abi.REDUCE_TF_SUM),
            "TS_sum": lib.reduce(h, _abi.REDUCE_TS_SUM), "FIX_sum": lib.reduce(h, _abi.REDUCE_FIX_SUM)}


def run(shape, steps, precision, collision, log):
    from . import _abi, workloads
    lib = _abi.load_library()
    t0 = time.perf_counter()
    lb = workloads.coupled3d(lib=lib, nx=shape[0], ny=shape[1], nz=shape[2], precision=precision, collision=collision)
    lb.init_simulation()
    log(f"[precision] {precision}: built in {time.perf_counter() - t0:.1f}s, running {steps} steps")
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        n = min(250, steps - done)
        lb.run(n)
        lb._lib.sync(lb._handle)
        done += n
        log(f"[precision] {precision}: {done}/{steps}")
    wall = time.perf_counter() - t0
    mlups = shape[0] * shape[1] * shape[2] * steps / wall / 1e6
    f = _fields(lb)
    bal = _balances(lb)
    lb._lib.destroy(lb._handle)
    lb._handle = None
    return f, bal, mlups


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", default="512,256,256")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--precisions", default="fp16,fp32")
    ap.add_argument("--collision", default="mrt", choices=["mrt", "bgk"])
    args = ap.parse_args(argv)
    shape = tuple(int(v) for v in args.shape.split(","))

    def log(m):
        print(m, file=sys.stderr, flush=True)

    ref, ref_bal, ref_mlups = run(shape, args.steps, "fp64", args.collision, log)
    for prec in args.precisions.split(","):
        f, bal, mlups = run(shape, args.steps, prec, args.collision, log)
        err = {}
        for k, b in ref.items():
            a = f[k].astype(np.float64)
            nb = float(np.linalg.norm(b))
            err[k] = float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))
        print(json.dumps({
            "workload": f"coupled3d {shape[0]}x{shape[1]}x{shape[2]} {args.collision}", "steps": args.steps,
            "storage": prec, "rel_l2_vs_fp64": err,
            "balances": {k: {"fp64": ref_bal[k], prec: bal[k], "rel_diff": (bal[k] - ref_bal[k]) / ref_bal[k]
                             if ref_bal[k] else bal[k]} for k in ref_bal},
            "mlups": {"fp64": round(ref_mlups, 1), prec: round(mlups, 1)},
        }), flush=True)


if __name__ == "__main__":
    main()
