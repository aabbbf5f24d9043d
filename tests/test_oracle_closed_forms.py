"""Closed-form checks of the oracle, independent of the reference binaries.

They pin what the golden vectors leave open (the 8th-order Holberg operator
of ADE8:263-266, which has no leapfrog reference trajectory) and would catch
plausible restatement mistakes in the rest: a wrong stencil index or sign, a
dropped coefficient, a transposed axis, a wrong profile exponent.
"""
import math

import numpy as np
import pytest

import oracle_lib as orc

HOLBERG = (1.231666, -1.041182e-1, 2.063707e-2, -3.570998e-3)


def staggered_symbol(order, k, delta):
    """Exact response of the staggered first-derivative operator to
    exp(i k x): D = (2/delta) * sum_m c_m sin((2m-1) k delta / 2)."""
    if order == 2:
        c = (1.0,)
    elif order == 4:
        c = (27.0 / 24.0, -1.0 / 24.0)
    else:
        c = HOLBERG
    return sum(cm * 2.0 * math.sin((2 * m + 1) * k * delta / 2.0)
               for m, cm in enumerate(c)) / delta


def _one_step(order, nx, ny, vx0=None, vy0=None, dx=5.0, dt=1e-4, lam=2.0e9, mu=1.0e9):
    zero_prof = lambda n: np.stack([np.zeros(n), np.ones(n), np.ones(n),  # noqa: E731
                                    np.zeros(n), np.ones(n), np.ones(n)])
    state = np.zeros((13, ny, nx))
    if vx0 is not None:
        state[0] = vx0
    if vy0 is not None:
        state[1] = vy0
    one = np.ones((ny, nx))
    return orc.run2d(nx, ny, order, dx, dx, dt, 1, lam * one, mu * one, 2000.0 * one,
                     zero_prof(nx), zero_prof(ny), 0, 0, None, None, [], [], state=state,
                     ewin=None)


@pytest.mark.parametrize("order", [2, 4, 8])
def test_stencil_dispersion_along_x_and_y(order):
    """One stress step from vx = sin(k x) (and vy = sin(k y)) without PML:
    sigma_xx(i+1/2) - 0 = dt (lambda+2mu) D vx exactly (up to rounding) at
    interior points, with D the closed-form staggered symbol; the y
    operator must give the same numbers on the transposed problem."""
    nx, ny, dx, dt, lam, mu = 96, 64, 5.0, 1e-4, 2.0e9, 1.0e9
    k = 2 * math.pi / (12 * dx)
    x = dx * np.arange(nx)
    vx0 = np.tile(np.sin(k * x), (ny, 1))
    o = _one_step(order, nx, ny, vx0=vx0, dx=dx, dt=dt, lam=lam, mu=mu)
    sym = staggered_symbol(order, k, dx)
    # derivative of sin(k x) at x + dx/2 is k cos(k (x + dx/2)); the discrete
    # operator gives sym * cos(k (x + dx/2))
    expect = dt * (lam + 2 * mu) * sym * np.cos(k * (x + dx / 2))
    r = order // 2
    sl = slice(r + 2, nx - r - 2)
    np.testing.assert_allclose(o["sxx"][5, sl], expect[sl], rtol=1e-9, atol=1e-3)
    np.testing.assert_allclose(o["syy"][5, sl], lam / (lam + 2 * mu) * expect[sl], rtol=1e-9,
                               atol=1e-3)
    # y direction: vy = sin(k y): sigma_yy(j) = dt (lambda+2mu) D vy at y - dy/2
    y = dx * np.arange(ny)
    vy0 = np.tile(np.sin(k * y)[:, None], (1, nx))
    o2 = _one_step(order, nx, ny, vy0=vy0, dx=dx, dt=dt, lam=lam, mu=mu)
    expect_y = dt * (lam + 2 * mu) * sym * np.cos(k * (y - dx / 2))
    sly = slice(r + 2, ny - r - 2)
    np.testing.assert_allclose(o2["syy"][sly, 7], expect_y[sly], rtol=1e-9, atol=1e-3)


def test_second_and_fourth_order_exact_on_polynomials():
    """2nd order is exact on quadratics, 4th order on quartics (staggered,
    centred differences): the one-step stress equals dt (lambda+2mu) f'."""
    nx, ny, dx, dt, lam, mu = 40, 16, 1.0, 1.0, 2.0, 1.0
    x = np.arange(nx, dtype=float)
    for order, poly in ((2, lambda t: 0.5 * t ** 2 - 3 * t), (4, lambda t: t ** 4 / 24 - t)):
        d = {2: lambda t: t - 3, 4: lambda t: t ** 3 / 6 - 1}[order]
        vx0 = np.tile(poly(x), (ny, 1))
        o = _one_step(order, nx, ny, vx0=vx0, dx=dx, dt=dt, lam=lam, mu=mu)
        expect = dt * (lam + 2 * mu) * d(x + 0.5)
        r = order // 2
        sl = slice(r, nx - r - 1)
        np.testing.assert_allclose(o["sxx"][4, sl], expect[sl], rtol=1e-12, atol=1e-9)


def test_profile_closed_form():
    """a-1 formulas at a few points computed independently (2D2:290-394)."""
    n, dx, dt, npml, cp, f0 = 60, 10.0, 1e-3, 10, 3000.0, 8.0
    p = orc.profile(n, dx, dt, npml, cp, f0)
    L = npml * dx
    d0 = -3.0 * cp * math.log(1e-3) / (2 * L)
    amax = math.pi * f0
    for i in (0, 3, 9, 10, 30, 52, 59):  # 0-based
        x = i * dx
        for half, (a, b, K) in ((False, (p[0], p[1], p[2])), (True, (p[3], p[4], p[5]))):
            xv = x + (dx / 2 if half else 0.0)
            dist = max(L - xv, xv - ((n - 1) * dx - L))
            if dist >= 0:
                xn = dist / L
                d, al = d0 * xn ** 2, max(amax * (1 - xn), 0.0)
            else:
                d, al = 0.0, 0.0
            bb = math.exp(-(d + al) * dt)
            aa = d * (bb - 1) / (d + al) if abs(d) > 1e-6 else 0.0
            assert K[i] == 1.0
            assert b[i] == pytest.approx(bb, rel=1e-14)
            assert a[i] == pytest.approx(aa, rel=1e-13, abs=1e-300)


def test_stf_closed_form():
    """a-6: first derivative of a Gaussian and Ricker, split at 135 degrees."""
    nst, dt, f0 = 400, 1e-3, 10.0
    t0, fac = 1.2 / f0, 1e7
    t = dt * np.arange(nst)
    a = math.pi ** 2 * f0 ** 2
    g1 = -fac * 2 * a * (t - t0) * np.exp(-a * (t - t0) ** 2)
    rk = fac * (1 - 2 * a * (t - t0) ** 2) * np.exp(-a * (t - t0) ** 2)
    ang = math.radians(135.0)
    fx, fy = orc.stf(nst, dt, f0, t0, fac, 135.0, 1)
    np.testing.assert_allclose(fx, math.sin(ang) * g1, rtol=1e-12, atol=1e-6)
    np.testing.assert_allclose(fy, math.cos(ang) * g1, rtol=1e-12, atol=1e-6)
    fx, fy = orc.stf(nst, dt, f0, t0, fac, 135.0, 2)
    np.testing.assert_allclose(fx, math.sin(ang) * rk, rtol=1e-12, atol=1e-6)


@pytest.mark.parametrize("order", [2, 4, 8])
def test_rigid_box_conserves_energy(order):
    """Without absorbing layers the discrete scheme conserves energy up to
    O(dt^2) oscillation: after the source stops, the total energy stays
    constant within a few parts in 1e3 (a sign error in any term breaks it)."""
    nx = ny = 64
    dx, dt = 10.0, 4e-4
    cp = 3000.0
    cs = cp / math.sqrt(3)
    rho = 2000.0 * np.ones((ny, nx))
    mu = rho * cs * cs
    lam = rho * (cp * cp - 2 * cs * cs)
    zero_prof = lambda n: np.stack([np.zeros(n), np.ones(n), np.ones(n),  # noqa: E731
                                    np.zeros(n), np.ones(n), np.ones(n)])
    nst = 600
    fx, fy = orc.stf(nst, dt, 20.0, 0.06, 1e7, 135.0, 1)
    fx[300:] = 0
    fy[300:] = 0
    o = orc.run2d(nx, ny, order, dx, dx, dt, nst, lam, mu, rho, zero_prof(nx), zero_prof(ny),
                  32, 32, fx, fy, [], [], ewin=(1, nx, 1, ny))
    e = o["ekin"] + o["epot"]
    tail = e[320:]
    assert tail.mean() > 0
    assert (tail.max() - tail.min()) / tail.mean() < 5e-3
