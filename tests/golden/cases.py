"""Parameter sets of the golden cases, written the way the reference programs
declare them (decimal literals evaluated in the program's precision).

Each case mirrors one variant built by oracle/build_ref.py.  ``values(case,
prec)`` evaluates the Fortran ``parameter`` expressions of the reference
(2D2:138-186, 2D4:132-187, 3D:124-196) in float64 or float32, rounding every
literal directly from its decimal text (as a Fortran compiler does) and every
operation in the target precision, in the Fortran evaluation order.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def literal(text: str, prec: str):
    """Correctly rounded value of a decimal literal in float64 or float32."""
    if prec == "f64":
        return np.float64(float(Fraction(text)))
    q = Fraction(text)
    x = np.float32(float(q))
    lo = np.nextafter(x, np.float32(-np.inf))
    hi = np.nextafter(x, np.float32(np.inf))

    def key(c):
        return (abs(Fraction(float(c)) - q), int(np.array(c, np.float32).view(np.uint32)) & 1)

    return min((lo, x, hi), key=key)


# Literal texts and integer parameters of each case.  Derived parameters are
# computed in values() in the reference's order.
CASES = {
    # 2D2 as shipped: 2D2:138-186
    "2d2_shipped": dict(order=2, NX=101, NY=641, DELTAX="10.", NPML=10, cp="3300.",
                        density="2800.", NSTEP=2000, DELTAT="2.e-3", f0="7.", factor="1.e7",
                        ANGLE="135.", NREC=2, src="gauss1", xdeb=None, model=False,
                        yorigin_n=False),
    # 2D4 as shipped: 2D4:132-187 (DELTAT = 2.e-3/2, NSTEP = 4000)
    "2d4_shipped": dict(order=4, NX=101, NY=641, DELTAX="10.", NPML=10, cp="3300.",
                        density="2800.", NSTEP=4000, DELTAT=("2.e-3", 2), f0="7.",
                        factor="1.e7", ANGLE="135.", NREC=2, src="gauss1", xdeb=None,
                        model=False, yorigin_n=True),
    # reduced heterogeneous layered cases (oracle/build_ref.py _LAYERED_2D)
    "2d2_layered": dict(order=2, NX=128, NY=96, DELTAX="5.", NPML=10, cp="5500.",
                        density="2800.", NSTEP=600, DELTAT="3.2e-4", f0="10.", factor="1.e7",
                        ANGLE="135.", NREC=8, src="ricker", ISOURCE=64, JSOURCE=48,
                        xdeb="50.", ydeb="425.", xfin="585.", yfin="425.", model=True,
                        yorigin_n=False, seed=20240611),
    "2d4_layered": dict(order=4, NX=128, NY=96, DELTAX="5.", NPML=10, cp="5500.",
                        density="2800.", NSTEP=600, DELTAT="3.2e-4", f0="10.", factor="1.e7",
                        ANGLE="135.", NREC=8, src="ricker", ISOURCE=64, JSOURCE=48,
                        xdeb="50.", ydeb="425.", xfin="585.", yfin="425.", model=True,
                        yorigin_n=True, seed=20240612),
    # reduced 3D (oracle/build_ref.py _SMALL_3D), 2 MPI ranks
    "3d_small": dict(order=2, NX=40, NY=44, NZ=36, DELTAX="10.", NPML=8, cp="3300.",
                     density="2800.", NSTEP=300, DELTAT="1.6e-3", f0="7.", factor="1.e7",
                     ANGLE="135.", NREC=2, src="gauss1", xdeb="-100", ydeb="300.",
                     xfin="0", yfin="100.", model=False, nproc=2),
}

# Cases without reference fixtures, checked against the oracle only.
# 3d_wide: 160 points along x (2.5 tiles of 64), so that some tiles touch no
# absorbing slab; the source lies in one of them.
OTHER_CASES = {
    "3d_wide": dict(CASES["3d_small"], NX=160, NY=48, NZ=40, NSTEP=200, ISOURCE=91,
                    JSOURCE=33),
    # the 3D benchmark medium at test size (layered along z, +-5% smoothed
    # perturbation, Courant 0.5): inputs from tests/layered_media.py only
    "3d_layered": dict(order=2, NX=96, NY=80, NZ=72, NPML=10, NSTEP=240),
}
# 3D 4th order: the elastic limit of the 3D viscoelastic program
# (oracle/build_ref.py _SMALL_3DV; 3D-visco:152-222), 2 MPI ranks, float64
# reference only.  Profiles and receiver indices are the reference's own
# (K_MAX_PML = 7, its own d0 and receiver rule), read from the fixture.
CASES_3D4 = {
    "3d4_small": dict(order=4, NX=40, NY=60, NZ=40, DELTAX="4.", NPML=8, cp="3000.",
                      cs="2000.", density="2000.", NSTEP=400, DELTAT="4.e-4", f0="18.",
                      factor="1.e7", ANGLE="0.", NREC=3, src="gauss1", xdeb=None, model=False,
                      ISOURCE=8 + 20, JSOURCE=60 // 5 + 1, nproc=2, precs=("f64",)),
}
ALL_CASES = {**CASES, **OTHER_CASES, **CASES_3D4}

VARIANT_OF = {
    "2d2_shipped": "ref2d2_shipped", "2d4_shipped": "ref2d4_shipped",
    "2d2_layered": "ref2d2_layered", "2d4_layered": "ref2d4_layered",
    "3d_small": "ref3d_small", "3d4_small": "ref3dv4_elastic",
}


def values(name: str, prec: str) -> dict:
    """Evaluate the case's parameters in precision ``prec`` ('f64' or 'f32')."""
    c = ALL_CASES[name]
    L = lambda s: literal(s, prec)  # noqa: E731
    R = np.float64 if prec == "f64" else np.float32
    v = dict(c)
    v["prec"] = prec
    v["DELTAX"] = L(c["DELTAX"])
    v["DELTAY"] = v["DELTAX"]
    v["DELTAZ"] = v["DELTAX"]
    if isinstance(c["DELTAT"], tuple):
        v["DELTAT"] = L(c["DELTAT"][0]) / R(c["DELTAT"][1])
    else:
        v["DELTAT"] = L(c["DELTAT"])
    v["cp"] = L(c["cp"])
    v["cs"] = L(c["cs"]) if "cs" in c else v["cp"] / L("1.732")
    v["density"] = L(c["density"])
    v["f0"] = L(c["f0"])
    v["t0"] = L("1.20") / v["f0"]
    v["factor"] = L(c["factor"])
    v["ANGLE"] = L(c["ANGLE"])
    nx, ny, npml = c["NX"], c["NY"], c["NPML"]
    v["ISOURCE"] = c.get("ISOURCE", nx - 2 * npml - 1)
    v["JSOURCE"] = c.get("JSOURCE", 2 * ny // 3 + 1)
    xsource = R(v["ISOURCE"] - 1) * v["DELTAX"]
    ysource = R(v["JSOURCE"] - 1) * v["DELTAY"]
    if c["xdeb"] is None:  # 2D2:180-183
        v["xdeb"] = xsource - L("100.")
        v["ydeb"] = L("2300.")
        v["xfin"] = xsource
        v["yfin"] = L("300.")
    elif name in ("3d_small", "3d_wide"):  # xdeb = xsource - 100, xfin = xsource
        v["xdeb"] = xsource - L("100.")
        v["ydeb"] = L(c["ydeb"])
        v["xfin"] = xsource
        v["yfin"] = L(c["yfin"])
    else:
        for k in ("xdeb", "ydeb", "xfin", "yfin"):
            v[k] = L(c[k])
    # homogeneous medium: 2D2:468-474, 3D:142-145
    d, cp, cs = v["density"], v["cp"], v["cs"]
    v["mu"] = d * cs * cs
    v["lambda"] = d * (cp * cp - R(2) * cs * cs)
    v["lambdaplustwomu"] = d * cp * cp
    return v


def layered_model(seed: int, nx: int, ny: int, prec: str, nlayers: int = 12):
    """Seeded horizontal layers along y (depth): cp ~ U[1500, 5500],
    cs = cp/sqrt(3), rho = 310 cp^0.25.  Returns rho, mu, lambda as (ny, nx)."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.integers(1, ny, size=nlayers - 1))
    cps = rng.uniform(1500.0, 5500.0, size=nlayers)
    layer = np.searchsorted(cuts, np.arange(ny), side="right")
    cp = cps[layer][:, None] * np.ones((1, nx))
    cs = cp / np.sqrt(3.0)
    rho = 310.0 * cp ** 0.25
    mu = rho * cs * cs
    lam = rho * (cp * cp - 2.0 * cs * cs)
    dt = np.float64 if prec == "f64" else np.float32
    return rho.astype(dt), mu.astype(dt), lam.astype(dt)
