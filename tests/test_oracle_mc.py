"""Pinning the Monte Carlo oracle without the reference (whose MEX source
needs MATLAB's mex.h and mx libraries, absent here): closed forms and exact
invariants of tests/closed_forms.py, each chosen so that a plausible
mistake in the function it covers (a dropped term, a wrong sign, index or
normalisation, a swapped axis) fails it:

  launch      pencil (beer_lambert), isotropic point (point_source) and line
              (line_source), plane wave with and without the /25 of an
              unbounded launch (planewave), radial top-hat / Gaussian focal
              profiles and top-hat / Gaussian / Lambertian angular profiles,
              X/Y profiles with psi, LG01 (Gamma(2) radius law), the 3D
              source (source3d: CDF sampling and its power normalisation)
  transport   Beer-Lambert, HG single-scatter histogram, tabulated phase
              function, Fresnel at normal and oblique incidence with Snell
              refraction, total internal reflection, energy balance
  boundaries  cyclic wrap, the too-old kill, escapes per face
  tallies     fiber/lens collectors (collected fraction, imaging, time bins),
              deposition criteria (immediate and end-of-life), example paths
  Lambert W   the identity W e^W = x and the reference's own lambert.c
              (compiled from its source into oracle/_ref where present)
"""
import math

import numpy as np
import pytest

import closed_forms as CF
import models
import oracle_lib as OL


def oracle_run(m, seed):
    return OL.oracle_mc(m, seed=seed)


CHECKS = ["beer_lambert", "fresnel_normal", "tir", "energy", "point_source", "planewave", "radial_focal_profiles",
          "radial_angular_profiles", "xy_profiles", "lg01", "cyclic_wrap", "too_old", "fresnel_oblique", "collectors",
          "unscattered_criteria", "example_paths", "source3d", "line_source"]


@pytest.mark.parametrize("check", CHECKS)
def test_oracle_closed_forms(check):
    getattr(CF, check)(oracle_run)


@pytest.mark.slow
def test_oracle_tabulated_phase():
    CF.tabulated_phase(oracle_run)


def test_lambert_wm1_identity():
    """W e^W = x on the lower branch (W <= -1) over [-1/e, 0)."""
    lib = OL.load_oracle()
    import ctypes as C

    lib.oracle_lambert_wm1.restype = C.c_double
    lib.oracle_lambert_wm1.argtypes = [C.c_double]
    xs = -np.exp(-1.0) * np.concatenate([np.linspace(1e-12, 1, 2001)[::-1], [1.0], 10.0 ** -np.arange(1, 15)])
    for x in xs:
        w = lib.oracle_lambert_wm1(float(x))
        assert w <= -1.0 + 1e-7
        res = w * math.exp(w) - x
        if x + math.exp(-1.0) < 1e-6:  # at the branch point the residual itself is the well-conditioned measure
            assert abs(res) <= 1e-15, (x, w, res)
        else:  # one Newton step's correction: the error in W implied by the residual
            assert abs(res / (math.exp(w) * (1 + w))) <= 1e-10 * abs(w), (x, w, res)


@pytest.mark.skipif(not __import__("os").path.exists(OL.REF_LAMBERT_LIB), reason="oracle/_ref (the reference's lambert.c) not built")
def test_lambert_wm1_matches_reference_source():
    """The restatement against gsl_sf_lambert_Wm1 of the reference's
    lambert.c, compiled from its source: the reference stops at a series
    approximation within q = x + 1/e < 3e-3 of the branch point, so the two
    agree to that series' accuracy there and to rounding elsewhere."""
    import ctypes as C

    lib = OL.load_oracle()
    lib.oracle_lambert_wm1.restype = C.c_double
    lib.oracle_lambert_wm1.argtypes = [C.c_double]
    ref = C.CDLL(OL.REF_LAMBERT_LIB)
    ref.gsl_sf_lambert_Wm1.restype = C.c_double
    ref.gsl_sf_lambert_Wm1.argtypes = [C.c_double]
    rng = np.random.default_rng(5)
    xs = -np.exp(-1.0) * np.concatenate([rng.random(5000), 10.0 ** -rng.uniform(0, 14, 2000)])
    for x in xs:
        a, b = lib.oracle_lambert_wm1(float(x)), ref.gsl_sf_lambert_Wm1(float(x))
        q = x + math.exp(-1.0)
        tol = 1e-6 if q < 3e-3 else 1e-13
        assert abs(a - b) <= tol * abs(b), (x, a, b)


@pytest.mark.slow
def test_oracle_hg_single_scatter():
    CF.hg_single(oracle_run)


def test_oracle_launch_failure_aborts():
    """A source aimed outside the cuboid: every launch fails, the run aborts
    with zero photons (MCmatlablib.c:708-710, MCmatlab.c:745)."""
    m = models.cases()["pencil_bt1"]()
    m.MC.LS.xFocus = 5.0  # far outside the 0.1 cm cube
    r = OL.oracle_mc(m, seed=1)
    assert r["nPhotons"] == 0
