"""What the trained network gains over the method without it, on a small coupled case.

The surrogate is trained here, on the device, by the library's own forward/backward kernels
(lbm_mlp_grad, lbm_mlp_adam) on samples that the plain FP64 integrator produced from two
sphere packings; it is then run on a third packing against the plain path:
  - it keeps temperature and solid conversion within 1e-2 of the plain path;
  - it replaces the stiff integration in (nearly) every cell, and the chemistry operator's
    device time per step drops several-fold;
  - an untrained network of the same shape, with the same normalisation and checks, misses
    the conversion by far more than the trained one.
"""
import numpy as np
import pytest

from lbm_porouschem_amd import chem, workloads
from lbm_porouschem_amd.chem import surrogate as S

pytestmark = pytest.mark.gpu

SHAPE = (32, 20, 20)
STEPS = 60


def run(lib, mode, sur, seed):
    nx, ny, nz = SHAPE
    lb = workloads.pyrolysis12(lib=lib, nx=nx, ny=ny, nz=nz, seed=seed)
    lb.init_simulation()
    g = chem.Chem(lib)
    if sur is not None:
        g.set_surrogate(sur)
    lb.attach_chemistry(g, mode, rho_solid_ref=lb.chem_rho_solid_ref)
    b0 = lb.species[12].S.to_numpy().astype(np.float64).sum()
    lb.run(STEPS)
    lb.sync()
    b1 = lb.species[12].S.to_numpy().astype(np.float64).sum()
    st = lb.chemistry_stats(g)
    return lb.TF.get_physical_value(lb.TF.S.to_numpy()), lb.TS.get_physical_value(lb.TS.S.to_numpy()), 1 - b1 / b0, st


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_trained_surrogate_gains_over_plain_and_untrained(hip):
    parts = [S.collect(hip, shape=SHAPE, steps=STEPS, every=5, packing_seed=sd, log=lambda m: None)
             for sd in (2025, 2027)]
    X = np.concatenate([p[0] for p in parts], axis=1)
    Y = np.concatenate([p[1] for p in parts], axis=1)
    U = np.concatenate([p[3] for p in parts])
    g = chem.Chem(hip)
    sur, hist = S.train(g, X, Y, U, epochs=25, batch=4096, log=lambda m: None)
    assert hist[-1]["holdout_loss"] < 0.2 * hist[0]["holdout_loss"]

    tf_p, ts_p, conv_p, st_p = run(hip, chem.MODE_PLAIN, None, 2026)
    tf_s, ts_s, conv_s, st_s = run(hip, chem.MODE_SURROGATE, sur, 2026)
    print(f"plain: conversion {conv_p:.4e}, chemistry {st_p['ms'] / STEPS:.3f} ms/step; surrogate: conversion "
          f"{conv_s:.4e}, chemistry {st_s['ms'] / STEPS:.3f} ms/step, fallbacks {st_s['fallback_hull']} hull "
          f"{st_s['fallback_residual']} residual of {st_s['cells']}; TF {rel(tf_s, tf_p):.2e} TS {rel(ts_s, ts_p):.2e}")
    assert rel(tf_s, tf_p) <= 1e-2 and rel(ts_s, ts_p) <= 1e-2
    assert abs(conv_s - conv_p) <= 1e-2 * conv_p
    assert st_s["fallback_hull"] + st_s["fallback_residual"] <= 0.02 * st_s["cells"]
    assert st_s["ms"] * 3 <= st_p["ms"]

    untrained = dict(sur)
    untrained["params"] = S.init_params(seed=1)
    tf_u, _, conv_u, st_u = run(hip, chem.MODE_SURROGATE, untrained, 2026)
    print(f"untrained: conversion {conv_u:.4e}, TF {rel(tf_u, tf_p):.2e}, fallbacks "
          f"{st_u['fallback_hull'] + st_u['fallback_residual']}")
    assert abs(conv_u - conv_p) >= 10 * max(abs(conv_s - conv_p), 1e-3 * conv_p)
