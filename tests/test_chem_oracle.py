"""Pinning of the CPU oracle's chemistry operator (oracle/chem_oracle.c) without the
product library: library ODE solvers, chemical equilibrium by Gibbs minimisation,
conservation laws, closed forms of the pyrolysis and fall-off limits, and torch
autograd for the MLP gradient.

The reference has no such operator (its reaction source is the explicit Arrhenius
update of src/LBM/LBM3D/_chemical.py:204-270), so there are no reference vectors;
each check below is chosen so that a plausible slip (wrong sign of dG, wrong
pressure factor of K_c, a dropped third body, ln for log10 in the Troe form, a
transposed weight in backprop) fails it.
"""
import ctypes
import json
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.optimize import minimize

from lbm_porouschem_amd import chem

R = chem.R_GAS
P_ATM = 101325.0


@pytest.fixture(scope="module")
def ochem(oracle64):
    lib = oracle64.lib
    lib.lbm_chem_oracle_rhs.argtypes = [ctypes.POINTER(chem.ChemMech), ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_void_p]
    lib.lbm_chem_oracle_rhs.restype = ctypes.c_int
    return oracle64


def rhs_fn(lib, mech, Csol):
    def f(t, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(14)
        assert lib.lib.lbm_chem_oracle_rhs(ctypes.byref(mech), y.ctypes.data, Csol, out.ctypes.data) == 0
        return out
    return f


def table():
    return chem.load_mechanism_table()


def nasa(sp, T):
    """cp/R, h/RT, s/R of one species from its NASA7 record (numpy, independent of the C code)."""
    a = np.array(sp["nasa_low"] if T < sp["T_ranges"][1] else sp["nasa_high"])
    cp = a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
    s = a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
    return cp, h, s


def state(T, Y, p=P_ATM, B=0.0):
    M = np.array([s["M"] for s in table()["species"]])
    rho = p / (R * T * np.sum(Y / M))
    return np.concatenate([rho * Y / M, [T, B]])


def syngas_air(co=0.05, h2=0.01, h2o=0.0):
    names = chem.species_names()
    Y = chem.air_mass_fractions()
    Y[names.index("CO")] = co
    Y[names.index("H2")] = h2
    Y[names.index("H2O")] = h2o
    return Y / Y.sum()


def test_mechanism_table_is_the_h2_co_subset_of_gri30():
    t = table()
    assert [s["name"] for s in t["species"]] == ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "CO", "CO2",
                                                  "HCO", "N2"]
    assert len(t["reactions"]) == 40
    kinds = [r["kind"] for r in t["reactions"]]
    assert kinds.count(0) == 32 and kinds.count(1) == 6 and kinds.count(2) + kinds.count(3) == 2
    # every reaction conserves the elements
    E = np.array([s["elements"] for s in t["species"]])
    for r in t["reactions"]:
        d = np.zeros(4)
        for j, nu in r["products"]:
            d += nu * E[j]
        for j, nu in r["reactants"]:
            d -= nu * E[j]
        assert np.all(d == 0), r["equation"]


def test_rate_constant_units_by_hand():
    """O + H2 <=> H + OH in GRI-Mech 3.0: A = 3.87e4 cm^3/(mol s), b = 2.7, Ea = 6260 cal/mol."""
    r = [x for x in table()["reactions"] if x["equation"] == "O + H2 <=> H + OH"][0]
    T = 1500.0
    k_cgs = 3.87e4 * T ** 2.7 * math.exp(-6260.0 * 4.184 / (R * T))
    k_si = r["A"] * T ** r["b"] * math.exp(-r["Ea_R"] / T)
    assert k_si == pytest.approx(k_cgs * 1e-6, rel=1e-12)
    # a three-body reaction carries one more concentration: 2 O + M <=> O2 + M, A = 1.2e17 cm^6/(mol^2 s)
    r3 = table()["reactions"][0]
    assert r3["equation"] == "2 O + M <=> O2 + M" and r3["A"] == pytest.approx(1.2e17 * 1e-12)


def test_gas_rhs_conserves_elements_and_energy(ochem):
    mech = chem.syngas_mechanism()
    E = np.array([[mech.elements[k][e] for e in range(4)] for k in range(12)])
    f = rhs_fn(ochem, mech, 0.0)
    rng = np.random.default_rng(3)
    sp = table()["species"]
    for _ in range(5):
        Y = rng.uniform(0.0, 1.0, 12) ** 3
        Y /= Y.sum()
        T = rng.uniform(900, 2200)
        y = state(T, Y)
        dy = f(0, y)
        w = dy[:12]
        scale = np.max(np.abs(w)) * np.max(np.abs(E)) + 1e-30
        assert np.max(np.abs(E.T @ w)) <= 1e-12 * scale
        # constant-volume adiabatic: d/dt sum_k c_k u_k(T) = 0 with u = (h/RT - 1) R T, cv = (cp/R - 1) R
        u = np.array([(nasa(s, T)[1] - 1.0) * R * T for s in sp])
        cv = np.array([(nasa(s, T)[0] - 1.0) * R for s in sp])
        dU = u @ w + (y[:12] @ cv) * dy[12]
        assert abs(dU) <= 1e-10 * (np.abs(u) @ np.abs(w))


def gibbs_equilibrium(T, p, n_elem, names_sp):
    """Equilibrium mole numbers at fixed T, p by minimising G over element-conserving compositions."""
    sp = table()["species"]
    E = np.array([s["elements"] for s in sp], dtype=float)
    g0 = np.array([nasa(s, T)[1] - nasa(s, T)[2] for s in sp])  # g/RT at 1 atm

    def G(x):
        n = np.exp(x)
        ntot = n.sum()
        return float(n @ (g0 + np.log(n / ntot * p / P_ATM)))

    def grad(x):
        n = np.exp(x)
        ntot = n.sum()
        return n * (g0 + np.log(n / ntot * p / P_ATM))

    cons = [{"type": "eq", "fun": (lambda x, e=e: E[:, e] @ np.exp(x) - n_elem[e]),
             "jac": (lambda x, e=e: E[:, e] * np.exp(x))} for e in range(4) if n_elem[e] > 0]
    x0 = np.log(np.full(12, 1e-3) + 0.1 * (E @ n_elem > 0))
    res = minimize(G, x0, jac=grad, constraints=cons, method="SLSQP", options={"maxiter": 2000, "ftol": 1e-16})
    return np.exp(res.x)


def test_long_time_kinetics_reaches_gibbs_equilibrium(ochem):
    """Reverse rates from K_c: integrating the kinetics to long times at fixed T must end at the
    composition that minimises the Gibbs energy (independent of the rate expressions)."""
    mech = chem.syngas_mechanism()
    T = 2400.0
    Y = syngas_air(co=0.10, h2=0.02)
    y0 = state(T, Y)
    # hold T fixed: a huge solid heat capacity makes dT/dt ~ 0
    f = rhs_fn(ochem, mech, 1e30)
    sol = solve_ivp(f, (0, 0.5), y0, method="LSODA", rtol=1e-10, atol=1e-16)
    c_end = sol.y[:12, -1]
    E = np.array([[mech.elements[k][e] for e in range(4)] for k in range(12)])
    n_elem = E.T @ y0[:12]
    # equilibrium at the final pressure of the constant-volume reactor
    p_end = c_end.sum() * R * T
    n_eq = gibbs_equilibrium(T, p_end, n_elem, None)
    x_kin = c_end / c_end.sum()
    x_eq = n_eq / n_eq.sum()
    big = x_eq > 1e-4
    assert np.max(np.abs(x_kin[big] - x_eq[big]) / x_eq[big]) < 2e-3
    # detailed balance: at equilibrium the net rates vanish
    w = f(0, np.concatenate([c_end, [T, 0.0]]))[:12]
    assert np.max(np.abs(w)) < 1e-6 * np.max(c_end)


def single_reaction_mech(eq):
    t = table()
    rec = [r for r in t["reactions"] if r["equation"] == eq][0]
    t2 = dict(t)
    t2["reactions"] = [rec]
    return chem.syngas_mechanism(table=t2), rec


def falloff_rate_fn(ochem, eq, T):
    mech, rec = single_reaction_mech(eq)
    for k in range(3):
        mech.pyro_A[k] = 0.0
    f = rhs_fn(ochem, mech, 1e30)
    names = chem.species_names()
    reac = rec["reactants"]
    prod = rec["products"][0][0]

    def rate_at(p):
        """forward rate coefficient (product absent) and the third-body concentration"""
        Y = np.zeros(12)
        Y[names.index("N2")] = 1.0
        c = state(T, Y, p=p)
        cM = c[:12].sum()
        for j, nu in reac:
            c[j] = 1e-6 * cM
        c[prod] = 0.0
        cm_eff = sum(rec["efficiencies"][j] * c[j] for j in range(12))
        conc = np.prod([c[j] ** nu for j, nu in reac])
        return f(0, c)[prod] / conc, cm_eff

    kinf = rec["A"] * T ** rec["b"] * math.exp(-rec["Ea_R"] / T)
    A0, b0, E0 = rec["low"]
    k0 = A0 * T ** b0 * math.exp(-E0 / T)
    return rate_at, rec, kinf, k0


def test_lindemann_falloff_limits(ochem):
    """Lindemann: k -> k_inf at high pressure, k -> k_0 [M] at low pressure."""
    rate_at, rec, kinf, k0 = falloff_rate_fn(ochem, "O + CO (+M) <=> CO2 (+M)", 1500.0)
    assert rec["kind"] == 2
    khi, _ = rate_at(1e14)
    assert khi == pytest.approx(kinf, rel=1e-4)
    klo, M = rate_at(1e-3)
    assert klo == pytest.approx(k0 * M, rel=1e-4)


def test_troe_falloff_special_points(ochem):
    """Troe form (Gilbert, Luther & Troe 1983): log10 F = log10 Fc / (1 + f1^2), f1 = (l + c) / (n - 0.14 (l + c)),
    l = log10 Pr.  At l = -c, F = Fc exactly; where n = 0.14 (l + c), F = 1 exactly (f1 infinite)."""
    T = 1500.0
    rate_at, rec, kinf, k0 = falloff_rate_fn(ochem, "2 OH (+M) <=> H2O2 (+M)", T)
    assert rec["kind"] == 3
    a, T3, T1, T2 = rec["troe"]
    fc = (1 - a) * math.exp(-T / T3) + a * math.exp(-T / T1) + (math.exp(-T2 / T) if T2 else 0.0)
    c = -0.4 - 0.67 * math.log10(fc)
    n = 0.75 - 1.27 * math.log10(fc)
    _, M1 = rate_at(P_ATM)
    for pr_target, F in ((10 ** (-c), fc), (10 ** (n / 0.14 - c), 1.0)):
        p = P_ATM * pr_target * kinf / (k0 * M1)
        k, M2 = rate_at(p)
        pr = k0 * M2 / kinf
        assert pr == pytest.approx(pr_target, rel=1e-9)
        assert k == pytest.approx(kinf * pr / (1 + pr) * F, rel=1e-6)


def test_ros2_matches_radau(ochem, oracle64):
    """The operator's ROS2 integration against scipy's Radau (rtol 1e-12) on the same right-hand
    side, from mild to igniting cells: at the default tolerance (rtol 1e-6) the temperature and
    the major species agree to the tolerance level, and with the tolerance tightened the whole
    state converges onto the Radau solution (a wrong stage coefficient would stall that)."""
    T = np.array([900.0, 1050.0, 1150.0, 1300.0])
    B = np.array([2.0, 2.0, 1.0, 0.0])
    Y = np.stack([syngas_air()] * 4, axis=1)
    x = chem.cell_inputs(T, P_ATM, Y, B, 0.1)
    dt = 1e-5
    mech0 = chem.syngas_mechanism()
    M = np.array([mech0.molemass[k] for k in range(12)])
    refs = []
    for i in range(4):
        y0 = state(T[i], Y[:, i], B=B[i])
        f = rhs_fn(ochem, mech0, (B[i] + 0.1) * mech0.cs_solid)
        sol = solve_ivp(f, (0, dt), y0, method="Radau", rtol=1e-12, atol=1e-18)
        refs.append(sol.y[:, -1])
    worst = {}
    for rtol, atol in ((1e-6, 1e-9), (1e-9, 1e-13)):
        mech = chem.syngas_mechanism(rtol=rtol, atol_c=atol, atol_T=1e-9, atol_B=1e-12, max_steps=1000000)
        C = chem.Chem(oracle64, mech)
        out = np.zeros((14, 4))
        ns = np.zeros(4, dtype=np.int32)
        C.integrate(x, out, ns, dt=dt)
        assert np.all(ns > 0)
        errs = []
        for i in range(4):
            y0 = state(T[i], Y[:, i], B=B[i])
            y = np.concatenate([y0[:12] + out[:12, i] / M, [T[i] + out[13, i], B[i] + out[12, i]]])
            ref = refs[i]
            major = np.abs(ref[:12]) > 1e-2 * ref[:12].max()
            errs.append((np.max(np.abs(y[:12][major] - ref[:12][major]) / np.abs(ref[:12][major])),
                         abs(y[12] - ref[12]) / ref[12],
                         np.max(np.abs(y[:12] - ref[:12]) / (np.abs(ref[:12]) + 1e-9))))
        worst[rtol] = np.max(np.array(errs), axis=0)
    assert worst[1e-6][0] < 2e-5 and worst[1e-6][1] < 1e-6
    assert worst[1e-9][0] < 1e-8 and worst[1e-9][1] < 1e-9 and worst[1e-9][2] < 1e-4


def test_pyrolysis_closed_form(oracle64):
    """No gas reactions possible (pure N2) and T pinned by a huge solid heat capacity: the biomass
    decays as exp(-k dt) with k the sum of the three routes, and the element residual of the
    result is round-off."""
    mech = chem.syngas_mechanism(cs_solid=1e30)
    C = chem.Chem(oracle64, mech)
    names = chem.species_names()
    Y = np.zeros(12)
    Y[names.index("N2")] = 1.0
    T = 1100.0
    x = chem.cell_inputs(T, P_ATM, Y[:, None], 2.0, 0.0)
    out = np.zeros((14, 1))
    dt = 1e-3
    C.integrate(x, out, None, dt=dt)
    k = sum(r["A"] * math.exp(-r["Ea"] / (R * T)) for r in chem.PYRO_ROUTES)
    assert out[12, 0] == pytest.approx(-2.0 * (1 - math.exp(-k * dt)), rel=1e-5)
    assert C.residual(x[:, 0], out[:, 0]) < 1e-12
    # the char that mass conservation assigns equals the char routes' carbon
    ks = [r["A"] * math.exp(-r["Ea"] / (R * T)) for r in chem.PYRO_ROUTES]
    conv = -out[12, 0]
    char = conv * (0.5 * ks[1] + 1.0 * ks[2]) / sum(ks) * chem.CHAR_M / chem.BIOMASS_M
    assert -out[12, 0] - out[:12, 0].sum() == pytest.approx(char, rel=1e-4)


def test_residual_flags_element_errors(oracle64):
    mech = chem.syngas_mechanism()
    C = chem.Chem(oracle64, mech)
    x = chem.cell_inputs(1200.0, P_ATM, syngas_air()[:, None], 1.0, 0.0)
    out = np.zeros((14, 1))
    C.integrate(x, out, None, dt=1e-5)
    assert C.residual(x[:, 0], out[:, 0]) < 1e-12
    bad = out[:, 0].copy()
    bad[chem.species_names().index("H2")] += 1e-4  # creates hydrogen out of nothing
    assert C.residual(x[:, 0], bad) > 1e-5


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -3.0e-3, 65504.0], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even significand -> 1; 1 + 3 * 2^-8 -> 1 + 2^-6
    np.testing.assert_array_equal(bf16(x)[:3], np.array([1.0, 1.0, 1.015625], dtype=np.float32))
    assert abs(bf16(x)[3] + 3.0e-3) <= 3.0e-3 * 2 ** -8


def torch_mlp_loss(P, xn, yt):
    """Same forward as the operator (BF16-rounded weights and layer inputs as straight-through
    estimators) in float64 torch; gradients come from autograd."""
    import torch

    def r(t):
        return t + (torch.from_numpy(bf16(t.detach().numpy().astype(np.float32)).astype(np.float64)) - t).detach()

    H = chem.HIDDEN
    o = 0
    W1 = P[o:o + H * 16].view(H, 16); o += H * 16
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].view(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    W3 = P[o:o + 14 * H].view(14, H); o += 14 * H
    b3 = P[o:o + 14]
    x = torch.from_numpy(bf16(xn.T).astype(np.float64))
    h1 = r(torch.relu(x @ r(W1).T + b1))
    h2 = r(torch.relu(h1 @ r(W2).T + b2))
    y = h2 @ r(W3).T + b3
    return ((y - torch.from_numpy(yt.T.astype(np.float64))) ** 2).mean()


def test_mlp_gradient_matches_autograd(oracle64):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    n = 96
    P = (rng.standard_normal(chem.NPARAM) * 0.1).astype(np.float32)
    xn = rng.standard_normal((16, n)).astype(np.float32)
    yt = rng.standard_normal((14, n)).astype(np.float32)
    C = chem.Chem(oracle64)
    g = np.zeros(chem.NPARAM, dtype=np.float32)
    loss = C.mlp_grad(P, xn, yt, g)
    Pt = torch.tensor(P.astype(np.float64), requires_grad=True)
    L = torch_mlp_loss(Pt, xn, yt)
    L.backward()
    ref = Pt.grad.numpy()
    assert loss == pytest.approx(float(L), rel=1e-6)
    assert np.linalg.norm(g - ref) <= 1e-5 * np.linalg.norm(ref)
    # forward of the oracle agrees with the loss it reports
    yn = np.zeros((14, n), dtype=np.float32)
    C.mlp_forward(P, xn, yn)
    assert float(np.mean((yn.astype(np.float64) - yt) ** 2)) == pytest.approx(loss, rel=1e-5)


def test_adam_matches_torch(oracle64):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    P = rng.standard_normal(chem.NPARAM).astype(np.float32)
    C = chem.Chem(oracle64)
    m = np.zeros_like(P)
    v = np.zeros_like(P)
    Pt = torch.tensor(P.copy(), requires_grad=True)
    opt = torch.optim.Adam([Pt], lr=1e-3)
    for t in range(1, 4):
        g = rng.standard_normal(chem.NPARAM).astype(np.float32)
        C.mlp_adam(P, g, m, v, 1e-3, t)
        Pt.grad = torch.from_numpy(g.copy())
        opt.step()
    np.testing.assert_allclose(P, Pt.detach().numpy(), rtol=1e-5, atol=1e-7)


def test_surrogate_output_transforms_invert_training_targets(oracle64):
    """The decode of lbm_chem_surrogate (out_kind 0/1/2, out_mul) is the inverse of the training
    targets of chem.surrogate: with the last layer zero the network's output is out_mean, so
    out_mean = the z targets of one sample must reproduce that sample's integrator outputs."""
    from lbm_porouschem_amd.chem import surrogate as S
    rng = np.random.default_rng(11)
    n = 40
    X = np.zeros((16, n))
    X[0] = rng.uniform(900, 1100, n)
    X[1] = P_ATM * rng.uniform(0.99, 1.01, n)
    X[2:14] = rng.uniform(0, 0.1, (12, n))
    X[14] = np.where(np.arange(n) % 4 == 0, 0.0, rng.uniform(0.5, 2.0, n))
    X[15] = rng.uniform(0, 0.5, n)
    Y = rng.standard_normal((14, n)) * np.logspace(-12, -4, n)[None, :]   # changes over many decades
    U = -np.exp(rng.uniform(-30, -5, n))                                   # conversion per unit biomass
    Y[12] = U * X[14]
    # species 1 (H) and 4 (OH) as end-of-step partial densities, the rest by the default transforms
    kinds = list(S.OUT_KIND)
    kinds[1] = kinds[4] = 3
    # radicals that reset within the step, to ~1e-9 kg/m^3 (the end-state form is chosen only
    # for such species: its absolute precision is one FP32 ulp of the partial density held)
    Y[1] = -0.9 * S.gas_density(X) * X[3] + 1e-9
    Y[4] = -0.95 * S.gas_density(X) * X[6] + 2e-9
    Z, kinds, lin = S.training_targets(X, Y, U, out_kind=kinds)
    assert np.all(lin[:12] > 0) and lin[13] > 0
    P = np.zeros(chem.NPARAM, dtype=np.float32)   # W3 = 0, b3 = 0: normalised outputs are 0
    C = chem.Chem(oracle64)
    for i in range(0, n, 7):
        sur = {"params": P, "in_mean": np.zeros(16, np.float32), "in_scale": np.ones(16, np.float32),
               "out_mean": Z[:, i].astype(np.float32), "out_scale": np.ones(14, np.float32),
               "hull_lo": np.full(16, -1e30, np.float32), "hull_hi": np.full(16, 1e30, np.float32),
               "residual_tol": 1e300, "out_mul": np.array(S.OUT_MUL, np.int32),
               "out_kind": np.array(kinds, np.int32), "out_lin": lin}
        C.set_surrogate(sur)
        x = np.ascontiguousarray(X[:, i:i + 1].astype(np.float32))
        y = np.zeros((14, 1), dtype=np.float32)
        flags = np.zeros(1, dtype=np.int32)
        C.surrogate(x, y, flags, fallback=0)
        assert flags[0] == 0
        ref = Y[:, i].copy()
        # inputs pass through float32 on the way in, so the per-unit output scales with float32(x)
        ref[12] = U[i] * float(np.float32(X[14, i]))
        np.testing.assert_allclose(y[:, 0], ref, rtol=5e-6, atol=0)


def test_nan_inputs_fall_out_of_the_hull(oracle64):
    """The hull test is an in-range test, so a NaN input (which compares false both ways) is
    outside the hull instead of inside it, and a finite cell next to it is unaffected."""
    rng = np.random.default_rng(5)
    n = 12
    x = chem.cell_inputs(rng.uniform(950, 1050, n), P_ATM * np.ones(n), np.repeat(syngas_air()[:, None], n, 1),
                         np.ones(n), np.zeros(n)).astype(np.float32)
    x[0, 3] = np.nan
    x[14, 8] = np.nan
    C = chem.Chem(oracle64)
    C.set_surrogate({"params": np.zeros(chem.NPARAM, np.float32), "in_mean": np.zeros(16, np.float32),
                     "in_scale": np.ones(16, np.float32), "out_mean": np.zeros(14, np.float32),
                     "out_scale": np.ones(14, np.float32), "hull_lo": np.full(16, -1e30, np.float32),
                     "hull_hi": np.full(16, 1e30, np.float32), "residual_tol": 1e30})
    out = np.zeros((14, n), np.float32)
    flags = np.zeros(n, np.int32)
    C.surrogate(x, out, flags, dt=1e-5, fallback=0)
    assert flags[3] == 1 and flags[8] == 1
    assert (flags == 0).sum() == n - 2
