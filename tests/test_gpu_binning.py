"""GPU media binning (getDownsampledParamVals) through the C ABI: bit for bit
the hand-worked answers of the reference's algorithm and the oracle's
outputs, on small and large inputs, with ties spread over many workgroups,
rows that do not vary, more bins than distinct sets and labels past 255."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as OL
from test_oracle_binning import worked_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from mcmatlab_amd import _lib, api as A

    assert _lib.lib().mcm_device_count() >= 1
    return A


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32:
        return a.shape == b.shape and np.array_equal(a.view(np.uint32), np.asarray(b, np.float32).view(np.uint32))
    return np.array_equal(a, b)


@pytest.mark.parametrize("case", worked_cases(), ids=lambda c: c[0])
def test_gpu_binning_matches_worked_cases(api, case):
    name, pv, n2, out, bins = case
    print(f"[gpu-binning] worked case {name}", flush=True)
    o, b = api.getDownsampledParamVals(pv, n2)
    assert same(o, out)
    assert same(b, bins)


def random_case(kind, n1, rng):
    if kind == "distinct":
        return rng.random((3, n1))
    if kind == "ties":
        return rng.integers(0, 5, (3, n1))
    if kind == "two_rows":
        pv = rng.random((3, n1))
        pv[1] = 120.0
        return pv
    if kind == "thermal":  # (VHC, TC, 1) as getThermalMediaProperties.m:75-78 builds them
        return np.vstack([3.5 + 0.4 * rng.random(n1), 0.004 + 0.001 * rng.random(n1), np.ones(n1)])
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["distinct", "ties", "two_rows", "thermal"])
@pytest.mark.parametrize("n1,n2", [(1, 3), (257, 20), (5000, 64), (70001, 255), (300000, 40)])
def test_gpu_binning_matches_oracle(api, kind, n1, n2):
    print(f"[gpu-binning] {kind} n1={n1} n2={n2}", flush=True)
    rng = np.random.default_rng(n1 + n2)
    pv = random_case(kind, n1, rng).astype(np.float32)
    o, b = api.getDownsampledParamVals(pv, n2)
    ro, rb = OL.oracle_binning(pv, n2)
    assert same(o, ro)
    assert same(b, rb)


def test_gpu_binning_many_workgroups_of_ties(api):
    """2^21 sets on a 4 x 4 x 4 integer grid: every pass has maxima tied
    across thousands of workgroups; the first (lowest-index) one must win,
    as in the reference's sequential scan."""
    rng = np.random.default_rng(11)
    pv = rng.integers(0, 4, (3, 1 << 21)).astype(np.float32)
    o, b = api.getDownsampledParamVals(pv, 70)
    ro, rb = OL.oracle_binning(pv, 70)
    assert same(o, ro) and same(b, rb)
    # more bins than the 64 distinct sets: every set sits at distance 0 of a pick
    assert len({tuple(c) for c in o.T}) == 64


def test_gpu_binning_labels_past_255_wrap(api):
    rng = np.random.default_rng(5)
    pv = rng.random((3, 4000)).astype(np.float32)
    o, b = api.getDownsampledParamVals(pv, 300)
    ro, rb = OL.oracle_binning(pv, 300)
    assert same(o, ro) and same(b, rb)
    assert b.max() == 255 and (b < 45).any()


def test_gpu_binning_invalid_arguments(api):
    from mcmatlab_amd._lib import MCMError

    with pytest.raises(MCMError):
        api.getDownsampledParamVals(np.zeros((3, 5), np.float32), 0)
    with pytest.raises(MCMError):
        api.getDownsampledParamVals(np.zeros((3, 0), np.float32), 3)
    with pytest.raises(ValueError):
        api.getDownsampledParamVals(np.zeros((2, 5), np.float32), 3)
    bad = np.ones((3, 5), np.float32)
    bad[1, 3] = np.nan
    with pytest.raises(MCMError):
        api.getDownsampledParamVals(bad, 3)


def test_gpu_binning_twice_in_a_row(api):
    """Device state (the pass counter, the picks) starts clean on every call."""
    rng = np.random.default_rng(9)
    a = rng.random((3, 20000)).astype(np.float32)
    c = rng.integers(0, 3, (3, 777)).astype(np.float32)
    r1 = api.getDownsampledParamVals(a, 50)
    r2 = api.getDownsampledParamVals(c, 9)
    r3 = api.getDownsampledParamVals(a, 50)
    assert same(r1[0], r3[0]) and same(r1[1], r3[1])
    ro, rb = OL.oracle_binning(c, 9)
    assert same(r2[0], ro) and same(r2[1], rb)


def test_gpu_bin_medium_sorts_and_relabels(api):
    """The callers' sortrows step: sets ascending lexicographically (mua,
    then mus, then g), and every label still names the same set."""
    from mcmatlab_amd.media import bin_medium
    from mcmatlab_amd.workloads import thermochromic_params

    pv = thermochromic_params(n=40)
    raw, rb = api.getDownsampledParamVals(pv, 30)
    ds, lab = bin_medium(pv, 30)
    keys = [tuple(c) for c in ds.T]
    assert keys == sorted(keys)
    assert sorted(map(tuple, raw.T)) == keys
    assert np.array_equal(ds[:, lab - 1], raw[:, rb.astype(np.int64) - 1])
    with pytest.raises(ValueError):
        bin_medium(pv, 300)

