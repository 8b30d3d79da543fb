"""Pins the MRT collision of the CPU oracle (oracle/lbm_oracle.c, mrt_collide).

The reference is BGK only (LBM3D/_evolution.py:33), so the multiple-relaxation-time
collision has no reference counterpart; it is fixed by what the method itself
implies:

  BGK reduction     every rate = 1/tau reproduces the BGK + Guo step to round-off
                    (checks the basis inverse, the equilibrium and forcing moments)
  conservation      mass (closed periodic box with gray and solid cells) and momentum
                    (no forcing) are conserved exactly
  shear viscosity   a decaying shear wave follows nu = (1/s_nu - 1/2)/3 when only the
                    shear-moment rates are changed (a mislabelled basis row fails this)
  energy moments    the rates of the isotropic energy moments leave the shear decay
                    unchanged; the heat-flux rates move it only at O(k^2)
"""
import numpy as np
import pytest

from lbm_porouschem_amd.flag import BC_MODEL, FLUID_STATE_EQUATION, FORCE_TERM
from lbm_porouschem_amd.solver import LBM2DSolver, LBM3DSolver

# shear-stress moments of the two bases (d'Humieres 2002 / Lallemand & Luo 2000 order)
SHEAR = {3: [9, 11, 13, 14, 15], 2: [7, 8]}
# the shear wave u_x(y) is carried by p_xy
PXY = {3: 13, 2: 8}


def _solver(oracle, dim, n, collision, **kw):
    if dim == 2:
        return LBM2DSolver(n[0], n[1], 1, dx=1, dt=1, precision="fp64", lib=oracle, collision=collision, **kw)
    return LBM3DSolver(n[0], n[1], n[2], dx=1, dt=1, precision="fp64", lib=oracle, collision=collision, **kw)


def _porous_case(oracle, dim, collision, eos=FLUID_STATE_EQUATION.IDEAL_GAS, rates=None):
    n = (14, 12) if dim == 2 else (9, 8, 7)
    lb = _solver(oracle, dim, n, collision, isThermal=True, isPoro=True)
    rng = np.random.default_rng(11)
    shape = (lb.nx, lb.ny, lb.nz)
    r = rng.random(shape)
    s = np.where(r < 0.3, 0.45, 0.0) + np.where(r > 0.92, 1.0, 0.0)
    lb.add_solid(s)
    lb.EOS = eos
    lb.force_term_model = FORCE_TERM.GUO
    lb.set_viscosity(0.08)
    lb.set_poro_Darcy(s > 0, 0.2)
    lb.set_fluid_Prandtl(0.71)
    lb.add_rho_solid((s > 0) * 1.0)
    lb.boundary_condition_model = BC_MODEL.NONE
    if rates is not None:
        lb.set_mrt_rates(rates)
    lb.init_field(lb.rho, 1 + 0.05 * rng.random(shape))
    lb.init_field3(lb.v, 0.03 * rng.random(shape), 0.02 * rng.random(shape),
                   (0.01 * rng.random(shape)) if dim == 3 else 0.0 * rng.random(shape))
    lb.init_field(lb.TF.S, rng.random(shape))
    return lb, s


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("eos", ["ideal", "incompressible"])
def test_mrt_with_bgk_rates_equals_bgk(oracle64, dim, eos):
    e = FLUID_STATE_EQUATION.IDEAL_GAS if eos == "ideal" else FLUID_STATE_EQUATION.INCOMPRESSIBLE
    A, _ = _porous_case(oracle64, dim, "mrt", e)  # all rates <= 0: every moment relaxes with 1/tau
    B, _ = _porous_case(oracle64, dim, "bgk", e)
    A.init_simulation()
    B.init_simulation()
    A.run(25)
    B.run(25)
    for fa, fb in ((A.v, B.v), (A.rho, B.rho), (A.TF.S, B.TF.S)):
        a, b = fa.to_numpy(), fb.to_numpy()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    # and a non-trivial rate set differs (the MRT path is really taken)
    C, _ = _porous_case(oracle64, dim, "mrt", e)
    C.set_mrt_standard_rates()
    C.init_simulation()
    C.run(25)
    assert np.linalg.norm(C.v.to_numpy() - B.v.to_numpy()) > 1e-8 * np.linalg.norm(B.v.to_numpy())


@pytest.mark.parametrize("dim", [2, 3])
def test_mrt_conserves_mass_in_closed_box(oracle64, dim):
    lb, s = _porous_case(oracle64, dim, "mrt")
    lb.set_mrt_standard_rates()
    lb.init_simulation()
    fluid = s < 1
    m0 = lb.rho.to_numpy()[fluid].sum()
    lb.run(50)
    m1 = lb.rho.to_numpy()[fluid].sum()
    assert abs(m1 - m0) <= 1e-12 * m0


@pytest.mark.parametrize("dim", [2, 3])
def test_mrt_conserves_momentum_without_forcing(oracle64, dim):
    n = (12, 10) if dim == 2 else (8, 7, 6)
    lb = _solver(oracle64, dim, n, "mrt")
    lb.set_viscosity(0.05)
    lb.set_mrt_standard_rates()
    lb.boundary_condition_model = BC_MODEL.NONE
    rng = np.random.default_rng(3)
    shape = (lb.nx, lb.ny, lb.nz)
    lb.init_field(lb.rho, 1 + 0.05 * rng.random(shape))
    lb.init_field3(lb.v, 0.04 * rng.random(shape) - 0.02, 0.04 * rng.random(shape) - 0.02,
                   (0.04 * rng.random(shape) - 0.02) if dim == 3 else 0.0 * rng.random(shape))
    lb.init_simulation()

    def momentum():
        return (lb.rho.to_numpy()[..., None] * lb.v.to_numpy()).reshape(-1, 3).sum(axis=0)

    p0 = momentum()
    lb.run(40)
    p1 = momentum()
    assert np.abs(p1 - p0).max() <= 1e-12 * lb.nx * lb.ny * lb.nz


def _shear_amplitude(oracle, dim, rates, steps, nu=0.05, ny=32, U=1e-3):
    n = (4, ny) if dim == 2 else (4, ny, 4)
    lb = _solver(oracle, dim, n, "mrt")
    lb.set_viscosity(nu)
    lb.set_mrt_rates(rates)
    lb.boundary_condition_model = BC_MODEL.NONE
    shape = (lb.nx, lb.ny, lb.nz)
    y = np.arange(ny)
    prof = np.sin(2 * np.pi * y / ny)
    ux = np.broadcast_to(U * prof[None, :, None], shape).copy()
    lb.init_field(lb.rho, np.ones(shape))
    lb.init_field3(lb.v, ux, np.zeros(shape), np.zeros(shape))
    lb.init_simulation()
    amps = []
    for s in steps:
        lb.run(s - lb.tLattice)
        u = lb.v.to_numpy()[..., 0].mean(axis=(0, 2))
        amps.append(2 / ny * (u * prof).sum())
    return np.array(amps)


@pytest.mark.parametrize("dim", [2, 3])
def test_shear_wave_decay_follows_shear_rate(oracle64, dim):
    """u_x = U sin(k y) decays as exp(-nu k^2 t).  With the shear moments relaxed at
    s_nu (and every other rate fixed), nu = (1/s_nu - 1/2)/3, independent of the
    viscosity set for tau."""
    ny = 32
    k = 2 * np.pi / ny
    t1, t2 = 50, 450
    for s_nu in (1.0, 1.6):
        rates = [0.0] * 19
        for r in SHEAR[dim]:
            rates[r] = s_nu
        A = _shear_amplitude(oracle64, dim, rates, [t1, t2], ny=ny)
        nu_meas = -np.log(A[1] / A[0]) / (k * k * (t2 - t1))
        nu_expect = (1 / s_nu - 0.5) / 3
        assert abs(nu_meas - nu_expect) <= 0.01 * nu_expect, (s_nu, nu_meas, nu_expect)
    # the p_xy rate alone controls this wave: changing another shear moment's rate does not
    rates = [0.0] * 19
    rates[PXY[dim]] = 1.6
    A = _shear_amplitude(oracle64, dim, rates, [t1, t2], ny=ny)
    nu_meas = -np.log(A[1] / A[0]) / (k * k * (t2 - t1))
    assert abs(nu_meas - (1 / 1.6 - 0.5) / 3) <= 0.01 * (1 / 1.6 - 0.5) / 3


@pytest.mark.parametrize("dim", [2, 3])
def test_ghost_rates_leave_shear_decay_unchanged(oracle64, dim):
    """The isotropic energy moments (e, eps; in 3D also pi_xx, pi_ww) do not couple to a
    transverse wave at linear order, so their rates change its decay only at O(U^2).
    (The heat-flux rates s_q do enter at O(k^2) through the dispersion, so they are
    not part of this check.)"""
    steps = [100, 300]
    base = _shear_amplitude(oracle64, dim, [0.0] * 19, steps)
    ghosts = [0.0] * 19
    for r in ((1, 2, 10, 12) if dim == 3 else (1, 2)):
        ghosts[r] = 1.3
    other = _shear_amplitude(oracle64, dim, ghosts, steps)
    assert np.allclose(other, base, rtol=1e-5, atol=0)
    # a heat-flux rate is a different moment: its effect is small but visible
    qr = [0.0] * 19
    for r in ((4, 6, 8) if dim == 3 else (4, 6)):
        qr[r] = 1.3
    q = _shear_amplitude(oracle64, dim, qr, steps)
    assert np.allclose(q, base, rtol=2e-3, atol=0) and not np.allclose(q, base, rtol=1e-7, atol=0)
