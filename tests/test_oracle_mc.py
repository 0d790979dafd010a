"""Pinning the Monte Carlo oracle.

1. Bit-for-bit against the reference C (+MCmatlab/src/MCmatlab.c and
   MCmatlablib.c, compiled from source in oracle/_ref with a frozen clock
   that fixes the dSFMT seed) on the feature matrix of tests/models.py:
   every output array and photon count must be identical.  Skipped where the
   reference build is absent.
2. Closed forms (tests/closed_forms.py), which need no reference.
"""
import numpy as np
import pytest

import closed_forms as CF
import models
import oracle_lib as OL

ref_only = pytest.mark.skipif(not OL.ref_available(), reason="reference build (oracle/_ref) not present")


@pytest.fixture(scope="module")
def ref_lib():
    from mexdriver import MexLib

    return MexLib(OL.REF_MC_LIB)


@ref_only
@pytest.mark.parametrize("name", list(models.cases()))
def test_oracle_bitwise_equals_reference(ref_lib, name):
    from mexdriver import run_mc_mex

    m = models.cases()[name]()
    seed = 11
    r = run_mc_mex(ref_lib, m, clock_us=seed)
    o = OL.oracle_mc(m, seed=seed)
    compared = 0
    for k, a in r.items():
        if k == "simulationTime" or k not in o:
            continue
        a = np.asarray(a).ravel(order="F")
        b = np.asarray(o[k]).ravel(order="F")
        assert a.shape == b.shape, (k, a.shape, b.shape)
        assert np.array_equal(a, b, equal_nan=True), k
        compared += 1
    assert compared >= 4


@ref_only
def test_oracle_bitwise_other_seeds(ref_lib):
    from mexdriver import run_mc_mex

    m = models.cases()["fresnel_bt3"]()
    for seed in (0, 1, 123456):
        r = run_mc_mex(ref_lib, m, clock_us=seed)
        o = OL.oracle_mc(m, seed=seed)
        assert np.array_equal(np.asarray(r["NFR"]).ravel(order="F"), np.asarray(o["NFR"]).ravel(order="F"))


@ref_only
def test_reference_interrupt_returns_partial_normalized_results(ref_lib):
    """Ctrl+C: the reference stops launching and normalises what it has
    (MCmatlab.c:293-298); the oracle restates it."""
    from mexdriver import run_mc_mex

    m = models.cases()["pencil_bt1"]()
    r = run_mc_mex(ref_lib, m, clock_us=1, interrupt_after=50)
    assert 0 < r["nPhotons"] < 2000
    assert np.all(np.isfinite(r["NFR"]))


def oracle_run(m, seed):
    return OL.oracle_mc(m, seed=seed)


@pytest.mark.parametrize("check", ["beer_lambert", "fresnel_normal", "tir", "energy", "point_source"])
def test_oracle_closed_forms(check):
    getattr(CF, check)(oracle_run)


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
