"""The MEX gateways (mcmatlab_amd/mex/*.c), built against the project's
mx-API shim: they read the same model fields as the reference MEX files and
write the same outputs.  The CPU tests drive them without a GPU (the call
must end in a MATLAB error naming the missing device, after reading every
field); the GPU tests check their outputs equal the Python API's for the same
packet streams."""
import os

import numpy as np
import pytest

import models
from mexdriver import MexLib, run_heat_mex, run_mc_mex

HERE = os.path.dirname(os.path.abspath(__file__))
MC_MEX = os.path.join(HERE, "_build", "libMCmatlab_mex_shim.so")
HEAT_MEX = os.path.join(HERE, "_build", "libfiniteElementHeatPropagator_mex_shim.so")
BIN_MEX = os.path.join(HERE, "_build", "libgetDownsampledParamVals_mex_shim.so")
built = pytest.mark.skipif(not (os.path.exists(MC_MEX) and os.path.exists(HEAT_MEX)), reason="MEX gateways not built")


def _no_gpu():
    from mcmatlab_amd import _lib

    return _lib.lib().mcm_device_count() == 0


@built
def test_mc_gateway_reads_every_field_then_reports_missing_device():
    if not _no_gpu():
        pytest.skip("a device is visible")
    m = MexLib(MC_MEX)
    for name in ("pencil_bt1", "radial_custom", "xy_tophat_gauss", "lens_timebins", "source3d", "criteria_collected"):
        with pytest.raises(RuntimeError, match="MCmatlab:GPU"):
            run_mc_mex(m, models.cases()[name]())


@built
def test_mc_gateway_rejects_invalid_model():
    m = MexLib(MC_MEX)
    bad = models.cases()["pencil_bt1"]()
    bad.MC.boundaryType = 7
    with pytest.raises(RuntimeError, match="MCmatlab:invalidInput"):
        run_mc_mex(m, bad)


@built
def test_heat_gateway_reports_missing_device():
    if not _no_gpu():
        pytest.skip("a device is visible")
    from test_gpu_heat import heat_case

    m = MexLib(HEAT_MEX)
    T0, Om0, P = heat_case()
    with pytest.raises(RuntimeError, match="MCmatlab:GPU"):
        run_heat_mex(m, T0, Om0, P)


@built
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["fresnel_bt1", "lens_timebins", "wavelengths3", "paths"])
def test_mc_gateway_equals_api(name, monkeypatch):
    from mcmatlab_amd import api

    monkeypatch.setenv("MCM_MC_SEED", "77")
    m = models.cases()[name]()
    r = run_mc_mex(MexLib(MC_MEX), m)
    a = api.MCmatlab(m, 1, seed=77).MC
    np.testing.assert_allclose(np.asarray(r["NFR"]).ravel(order="F"), np.asarray(a.NFR).ravel(order="F"), rtol=1e-6)
    assert r["nPhotons"] == a.nPhotons
    for k in ("NI_zneg", "NI_zpos"):
        np.testing.assert_allclose(np.asarray(r[k]).ravel(order="F"), np.asarray(getattr(a, k)).ravel(order="F"), rtol=1e-6)
    if m.MC.useLightCollector:
        np.testing.assert_allclose(np.asarray(r["image"]).ravel(order="F"), np.asarray(a.LC.image).ravel(order="F"), rtol=1e-6)
    if m.MC.nExamplePaths:
        np.testing.assert_array_equal(r["examplePaths"], a.examplePaths)


@built
@pytest.mark.gpu
def test_heat_gateway_equals_api():
    from mcmatlab_amd import api
    from test_gpu_heat import heat_case

    T0, Om0, P = heat_case()
    r = run_heat_mex(MexLib(HEAT_MEX), T0, Om0, P)
    T, Om, sens, mx = api.finiteElementHeatPropagator(T0, Om0, P)
    np.testing.assert_array_equal(r["T"], T)
    np.testing.assert_array_equal(r["Omega"], Om)
    np.testing.assert_array_equal(r["sensorTemps"], sens)
    np.testing.assert_array_equal(r["maxMediaTemps"].ravel(), mx.ravel())


@built
@pytest.mark.gpu
def test_heat_gateway_precision_field():
    """P.precision = 'single' selects the FP32 arithmetic, 'double' the FP64
    default; anything else is an input error."""
    from mcmatlab_amd import api
    from test_gpu_heat import heat_case

    T0, Om0, P = heat_case()
    m = MexLib(HEAT_MEX)
    r32 = run_heat_mex(m, T0, Om0, dict(P, precision="single"))
    T32, Om32, _, _ = api.finiteElementHeatPropagator(T0, Om0, P, precision="fp32")
    np.testing.assert_array_equal(r32["T"], T32)
    np.testing.assert_array_equal(r32["Omega"], Om32)
    r64 = run_heat_mex(m, T0, Om0, dict(P, precision="double"))
    T64, _, _, _ = api.finiteElementHeatPropagator(T0, Om0, P)
    np.testing.assert_array_equal(r64["T"], T64)
    assert not np.array_equal(r32["T"], r64["T"])
    with pytest.raises(RuntimeError, match="invalidInput"):
        run_heat_mex(m, T0, Om0, dict(P, precision="half"))


def _bin_mex(m, pv, n2):
    from mexdriver import mxSINGLE

    a = m.num(np.asarray(pv, np.float32), mxSINGLE, shape=np.asarray(pv).shape)
    b = m.num(float(n2))
    try:
        outs = m.call(2, [a, b])
        res = m.array(outs[0]), m.array(outs[1])
        for o in outs:
            m.lib.mxs_destroy(o)
        return res
    finally:
        m.lib.mxs_destroy(a)
        m.lib.mxs_destroy(b)


@pytest.mark.skipif(not os.path.exists(BIN_MEX), reason="binning gateway not built")
def test_binning_gateway_checks_arguments_then_reports_missing_device():
    m = MexLib(BIN_MEX)
    with pytest.raises(RuntimeError, match="MCmatlab:invalidInput"):
        _bin_mex(m, np.zeros((2, 4), np.float32), 3)
    if not _no_gpu():
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError, match="MCmatlab:GPU"):
        _bin_mex(m, np.random.default_rng(0).random((3, 50)), 4)


@pytest.mark.skipif(not os.path.exists(BIN_MEX), reason="binning gateway not built")
@pytest.mark.gpu
def test_binning_gateway_equals_worked_cases():
    """The gateway called with the mxArrays the MATLAB callers pass returns
    the worked answers of the reference's algorithm with the reference's
    output classes and sizes."""
    from test_oracle_binning import worked_cases

    m = MexLib(BIN_MEX)
    for name, pv, n2, out, bins in worked_cases():
        o, b = _bin_mex(m, pv, n2)
        assert o.dtype == np.float32 and o.shape == (3, n2), name
        assert b.dtype == np.uint8 and b.shape == (1, pv.shape[1]), name
        assert np.array_equal(o.view(np.uint32), out.view(np.uint32)), name
        assert np.array_equal(b.reshape(-1), bins), name

