"""Learned replacement of the plain chemistry integrator: data from the plain path, training
on the device through the library's own forward/backward kernels, and the surrogate record
that lbm_chem_set_surrogate loads.

    X, Y, dt, U = collect(hip, shape=(96, 64, 64), steps=100, every=10)  # plain FP64 integrator
    sur, log = train(gchem, X, Y, U)                                       # lbm_mlp_grad + lbm_mlp_adam
    save(path, sur)

Torch appears here only as device memory for the C ABI's array arguments.
"""
import time

import numpy as np

from . import Chem, HIDDEN, MODE_PLAIN, NIN, NOUT, NPARAM


def cell_inputs_from_solver(lb, p0=101325.0, rho_solid_ref=2.0):
    """Operator inputs [16, n] of the solver's cells with solid < 1 (the cells the operator
    evaluates), built from the fields exactly as the attached operator builds them."""
    solid = lb.solid.to_numpy().ravel().astype(np.float64)
    live = solid < 1
    TF = lb.TF.S.to_numpy().ravel().astype(np.float64)
    TS = lb.TS.S.to_numpy().ravel().astype(np.float64)
    T = np.where(solid > 0, TS * lb.TS.v_scale + lb.TS.v_ref, TF * lb.TF.v_scale + lb.TF.v_ref)
    x = np.empty((NIN, int(live.sum())))
    x[0] = T[live]
    x[1] = p0 * lb.rho.to_numpy().ravel().astype(np.float64)[live]
    for q in range(12):
        x[2 + q] = lb.species[q].S.to_numpy().ravel().astype(np.float64)[live]
    x[14] = lb.species[12].S.to_numpy().ravel().astype(np.float64)[live] * rho_solid_ref
    x[15] = lb.species[13].S.to_numpy().ravel().astype(np.float64)[live] * rho_solid_ref
    return x


MAJOR = ("H2", "O2", "H2O", "CO", "CO2")


def jitter(x, rng):
    """A perturbed copy of logged cell states: T +-8 K, p +-2 %, trace species x U(0, 2), major
    species x U(0.9, 1.1), N2 closing the mass fractions, biomass x U(0.9, 1), char x U(0.8, 1.2).
    Training on these as well teaches the network the neighbourhood of the trajectories, which
    is where a coupled run with the surrogate's own small errors goes."""
    from . import species_names
    names = species_names()
    n = x.shape[1]
    xa = x.copy()
    xa[0] += rng.uniform(-8.0, 8.0, n)
    xa[1] *= rng.uniform(0.98, 1.02, n)
    for k, name in enumerate(names):
        if name == "N2":
            continue
        if name in MAJOR:
            xa[2 + k] = np.maximum(xa[2 + k] * rng.uniform(0.9, 1.1, n), 0.0)
        else:
            xa[2 + k] *= rng.uniform(0.0, 2.0, n)
    k2 = names.index("N2")
    xa[2 + k2] = 1.0 - (xa[2:14].sum(axis=0) - xa[2 + k2])
    xa[14] *= rng.uniform(0.9, 1.0, n)
    xa[15] *= rng.uniform(0.8, 1.2, n)
    return xa


# biomass density [kg/m^3] given to cells without biomass when their conversion rate per unit of
# biomass is measured (the rate depends on the cell's temperature and composition only)
B_PROBE = 1.0


def collect(lib, shape=(96, 64, 64), steps=300, every=20, max_samples=2_000_000, seed=0, precision="fp32",
            packing_seed=2025, jittered=True, log=print):
    """Run the coupled pyrolysis case with the plain operator attached and log (inputs, outputs)
    of the plain FP64 integrator every `every` steps, plus (jittered=True) the integrator's
    outputs on a perturbed copy of each snapshot.  Returns (X [16, n], Y [14, n], dt, U [n]):
    U is the biomass change per unit of biomass over the step, measured for cells without
    biomass by a second integration with B_PROBE of biomass."""
    import torch

    from .. import workloads
    rng = np.random.default_rng(seed + packing_seed)
    nx, ny, nz = shape
    lb = workloads.pyrolysis12(lib=lib, nx=nx, ny=ny, nz=nz, precision=precision, seed=packing_seed)
    lb.init_simulation()
    g = Chem(lib)
    lb.attach_chemistry(g, MODE_PLAIN, rho_solid_ref=lb.chem_rho_solid_ref)
    xs, ys, us = [], [], []
    t0 = time.perf_counter()

    def integrate(x):
        xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        yd = torch.zeros((NOUT, x.shape[1]), dtype=torch.float64, device="cuda")
        nd = torch.zeros(x.shape[1], dtype=torch.int32, device="cuda")
        g.integrate(xd, yd, nd, dt=lb.dt)
        return yd.cpu().numpy(), nd.cpu().numpy()

    for s in range(0, steps + 1, every):
        if s:
            lb.run(every)
        x0 = cell_inputs_from_solver(lb, rho_solid_ref=lb.chem_rho_solid_ref)
        for x in ((x0, jitter(x0, rng)) if jittered else (x0,)):
            y, ns = integrate(x)
            u = np.zeros(x.shape[1])
            has = x[14] > 0
            u[has] = y[12, has] / x[14, has]
            if (~has).any():
                xp = x[:, ~has].copy()
                xp[14] = B_PROBE
                u[~has] = integrate(xp)[0][12] / B_PROBE
            xs.append(x)
            ys.append(y)
            us.append(u)
        log(f"collect: packing {packing_seed} step {s}: {x0.shape[1]} cells, T {x0[0].min():.0f}-{x0[0].max():.0f} K, "
            f"integrator steps mean {ns.mean():.1f} max {ns.max()} ({time.perf_counter() - t0:.1f} s)")
    X = np.concatenate(xs, axis=1)
    Y = np.concatenate(ys, axis=1)
    U = np.concatenate(us)
    if X.shape[1] > max_samples:
        idx = np.random.default_rng(seed).choice(X.shape[1], max_samples, replace=False)
        X, Y, U = X[:, idx], Y[:, idx], U[idx]
    return X, Y, lb.dt, U


def init_params(seed=0):
    """He-normal weights for the ReLU layers, zero biases."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((16, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, NOUT)):
        P[o:o + fan_in * fan_out] = rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in)
        o += fan_in * fan_out + fan_out
    return P


def normalisation(X, Y, margin=0.02):
    """Per-feature z-score of inputs and outputs; hull = training-input bounding box widened
    by `margin` of its range.  An output that is constant over the data (the change of an
    inert species) gets scale 0: the surrogate then returns that constant exactly."""
    in_mean = X.mean(axis=1)
    in_sd = X.std(axis=1)
    in_sd[in_sd == 0] = 1.0
    out_mean = Y.mean(axis=1)
    out_sd = Y.std(axis=1)
    lo, hi = X.min(axis=1), X.max(axis=1)
    span = hi - lo
    return {
        "in_mean": in_mean.astype(np.float32), "in_scale": (1.0 / in_sd).astype(np.float32),
        "out_mean": out_mean.astype(np.float32), "out_scale": out_sd.astype(np.float32),
        "hull_lo": (lo - margin * span).astype(np.float32), "hull_hi": (hi + margin * span).astype(np.float32),
    }


# Output transforms (lbm_surrogate.out_kind):
#  - output 12 (biomass change) is learned per unit of biomass (input 14), as the logarithm of the
#    conversion per step (kind 1): first-order Arrhenius decay, a function of the temperature (and
#    weakly the composition) that spans decades, always negative, exactly zero without biomass;
#  - gas species that relax to a quasi-steady value within one step (their change cancels most of
#    what they held: median |d rho_k| / rho_k > FAST_RESET over the data) are learned as the
#    logarithm of their partial density at the end of the step (kind 3): learning the change would
#    ask the network to cancel the current value exactly, and its error then drives the species
#    negative or lets it drift by orders of magnitude;
#  - the other gas species changes and the temperature change are learned as asinh(y / s) (kind 2),
#    s = the median magnitude of the output over the data: linear in the many quiet cells, whose
#    tiny changes must stay tiny, logarithmic in the few reacting ones.
OUT_MUL = [-1] * 12 + [14, -1]
OUT_KIND = [2] * 12 + [1, 2]   # before the choice of kind-3 species (choose_kinds)
FAST_RESET = 0.5
LOG_FLOOR = 1e-20
RHO_FLOOR = 1e-30
LIN_FLOOR = 1e-30


def gas_density(X):
    """rho_g = p / (R T sum_k Y_k / M_k) of each sample [kg/m^3] (the decode's formula)."""
    from . import R_GAS, load_mechanism_table
    M = np.array([sp["M"] for sp in load_mechanism_table()["species"]])
    sym = (X[2:14] / M[:, None]).sum(axis=0)
    ok = (sym > 0) & (X[0] > 0)
    return np.where(ok, X[1] / (R_GAS * np.where(ok, X[0], 1.0) * np.where(ok, sym, 1.0)), 0.0)


def choose_kinds(X, Y, base=OUT_KIND):
    """out_kind per output: kind 3 for the gas species that reset within a step."""
    kinds = list(base)
    have = gas_density(X)[None, :] * X[2:14]
    for k in range(12):
        live = have[k] > 0
        if live.sum() and np.median(np.abs(Y[k, live]) / have[k, live]) > FAST_RESET:
            kinds[k] = 3
    return kinds


def per_unit(X, Y, U=None, out_mul=OUT_MUL):
    """Outputs predicted per unit of an input (U, when given, holds the per-unit biomass change of
    every sample; otherwise Y / X where X > 0 and 0 elsewhere)."""
    T = np.array(Y, dtype=np.float64, copy=True)
    for o, v in enumerate(out_mul):
        if v >= 0:
            if U is not None and o == 12:
                T[o] = U
            else:
                d = X[v]
                T[o] = np.divide(Y[o], d, out=np.zeros_like(Y[o]), where=d > 0)
    return T


def linear_scales(T, out_kind=OUT_KIND):
    """out_lin of the kind-2 outputs: the median magnitude of the nonzero values (1 elsewhere)."""
    lin = np.ones(len(out_kind), dtype=np.float64)
    for o, k in enumerate(out_kind):
        if k == 2:
            a = np.abs(T[o])
            a = a[a > 0]
            lin[o] = max(float(np.median(a)), LIN_FLOOR) if a.size else 1.0
    return lin.astype(np.float32)


def training_targets(X, Y, U=None, out_mul=OUT_MUL, out_kind=None, out_lin=None):
    """Network targets before normalisation (z of lbm_surrogate): per-unit outputs, then the
    inverse of each output's transform.  Returns (targets, out_kind, out_lin)."""
    if out_kind is None:
        out_kind = choose_kinds(X, Y)
    T = per_unit(X, Y, U, out_mul)
    if out_lin is None:
        out_lin = linear_scales(T, out_kind)
    rho_g = gas_density(X)
    for o, k in enumerate(out_kind):
        if k == 1:
            T[o] = np.log(np.maximum(-T[o], LOG_FLOOR))
        elif k == 2:
            T[o] = np.arcsinh(T[o] / np.float64(out_lin[o]))
        elif k == 3:
            T[o] = np.log(np.maximum(rho_g * X[2 + o] + T[o], RHO_FLOOR))
    return T, list(out_kind), out_lin


def normalise(X, Y, norm):
    Xn = ((X.astype(np.float32) - norm["in_mean"][:, None]) * norm["in_scale"][:, None]).astype(np.float32)
    sd = norm["out_scale"].astype(np.float64)
    Yn = ((Y - norm["out_mean"][:, None]) / np.where(sd > 0, sd, 1.0)[:, None]).astype(np.float32)
    return Xn, Yn


# loss weight (as a divisor of the output scale, so the weight is its square) of the outputs
# that the coupled run integrates over many steps: the biomass change and the temperature change
OUT_EMPHASIS = {12: 3.0, 13: 3.0}


def train(g, X, Y, U=None, epochs=40, batch=8192, lr=1e-3, decay=0.92, holdout=0.05, residual_tol=1e-3, seed=0,
          params=None, margin=0.5, emphasis=OUT_EMPHASIS, log=print):
    """Fit the 16-128-128-14 MLP on the device with the library's kernels (lbm_mlp_grad +
    lbm_mlp_adam); the learning rate decays by `decay` per epoch.  U: per-unit biomass change of
    every sample (collect()).  Returns (surrogate dict, history; history[0] is the untrained
    network)."""
    import torch
    rng = np.random.default_rng(seed)
    Y, out_kind, out_lin = training_targets(X, Y, U)
    n = X.shape[1]
    perm = rng.permutation(n)
    nh = max(int(n * holdout), 1)
    hold, fit = perm[:nh], perm[nh:]
    norm = normalisation(X[:, fit], Y[:, fit], margin=margin)
    for o, w in (emphasis or {}).items():
        norm["out_scale"][o] /= w
    Xn, Yn = normalise(X, Y, norm)
    Xd = torch.from_numpy(np.ascontiguousarray(Xn[:, fit])).cuda()
    Yd = torch.from_numpy(np.ascontiguousarray(Yn[:, fit])).cuda()
    Xh = torch.from_numpy(np.ascontiguousarray(Xn[:, hold])).cuda()
    Yh = torch.from_numpy(np.ascontiguousarray(Yn[:, hold])).cuda()
    P = torch.from_numpy(params if params is not None else init_params(seed)).cuda()
    G = torch.zeros_like(P)
    M = torch.zeros_like(P)
    V = torch.zeros_like(P)
    nf = Xd.shape[1]
    steps_per_epoch = max(nf // batch, 1)
    # epoch -1: the network before any update, the baseline the training is measured against
    hist = [{"epoch": -1, "train_loss": None, "holdout_loss": holdout_loss(g, P, Xh, Yh), "seconds": 0.0}]
    t = 0
    t0 = time.perf_counter()
    rate = lr
    for ep in range(epochs):
        order = torch.randperm(nf, device="cuda")
        tot = 0.0
        for b in range(steps_per_epoch):
            idx = order[b * batch:(b + 1) * batch]
            xb = Xd[:, idx].contiguous()
            yb = Yd[:, idx].contiguous()
            tot += g.mlp_grad(P, xb, yb, G)
            t += 1
            g.mlp_adam(P, G, M, V, rate, t)
        rate *= decay
        hl = holdout_loss(g, P, Xh, Yh)
        hist.append({"epoch": ep, "train_loss": tot / steps_per_epoch, "holdout_loss": hl,
                     "seconds": time.perf_counter() - t0})
        log(f"train: epoch {ep}: loss {tot / steps_per_epoch:.3e} holdout {hl:.3e} "
            f"({time.perf_counter() - t0:.1f} s)")
    sur = dict(norm)
    sur["params"] = P.cpu().numpy()
    sur["residual_tol"] = residual_tol
    sur["out_mul"] = np.array(OUT_MUL, dtype=np.int32)
    sur["out_kind"] = np.array(out_kind, dtype=np.int32)
    sur["out_lin"] = out_lin
    return sur, hist


def holdout_loss(g, P, Xh, Yh):
    import torch
    out = torch.zeros_like(Yh)
    g.mlp_forward(P, Xh, out)
    return float(((out - Yh) ** 2).mean())


def save(path, sur):
    np.savez(path, **{k: np.asarray(v) for k, v in sur.items()})


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: (z[k] if z[k].ndim else z[k].item()) for k in z.files}
