"""The media-binning oracle (oracle/binning_oracle.c) against the reference's
getDownsampledParamVals.c: the committed golden vectors of the reference
build, the reference build itself where oracle/_ref is built, and
brute-force properties of farthest-point sampling that a plausible mistake
(a wrong distance, a lost tie rule, labels off by one) breaks."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle_lib as OL

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "binning_vectors.npz")


def golden_cases():
    z = np.load(GOLDEN)
    keys = sorted({k.rsplit("_", 1)[0] for k in z.files}, key=lambda k: int(k.split("_")[0]))
    return [(k, z[k + "_pv"], int(z[k + "_n2"]), z[k + "_out"], z[k + "_bin"]) for k in keys]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_oracle_matches_reference_vectors(case):
    _, pv, n2, out, bins = case
    o, b = OL.oracle_binning(pv, n2)
    assert np.array_equal(o.view(np.uint32), out.view(np.uint32))
    assert np.array_equal(b, bins)


@pytest.mark.skipif(not os.path.exists(OL.REF_BINNING_LIB), reason="reference build (oracle/_ref) not present")
def test_oracle_matches_reference_build_random():
    rng = np.random.default_rng(7)
    for t in range(150):
        n1, n2 = int(rng.integers(1, 300)), int(rng.integers(1, 40))
        kind = t % 4
        if kind == 0:
            pv = rng.random((3, n1))
        elif kind == 1:
            pv = rng.integers(0, 4, (3, n1))
        elif kind == 2:
            pv = rng.integers(0, 3, (3, n1)) * 0.37
            pv[2] = 0.9
        else:
            pv = np.vstack([rng.random(n1), np.full(n1, 2.0), rng.integers(0, 2, n1)])
        pv = pv.astype(np.float32)
        o, b = OL.oracle_binning(pv, n2)
        r, rb = OL.ref_binning(pv, n2)
        assert np.array_equal(o.view(np.uint32), r.view(np.uint32)), (t, n1, n2)
        assert np.array_equal(b, rb), (t, n1, n2)


def _normalised(pv):
    pv = pv.astype(np.float32)
    rng_ = pv.max(axis=1) - pv.min(axis=1)
    rows = [r for r in range(3) if rng_[r] != 0]
    return np.array([pv[r] / rng_[r] for r in rows], np.float64).reshape(len(rows), -1)


def test_oracle_farthest_point_properties():
    """Distinct sets: the picks are distinct input sets, the first one is the
    set farthest from the mean and the second the set farthest from the
    first (up to float rounding of the distances), and every label names a
    nearest pick.  Later picks follow the reference's scan, whose running
    maximum starts at the previous pass's index, so they need not be the
    farthest set; the reference build and its vectors pin those."""
    rng = np.random.default_rng(3)
    pv = rng.random((3, 200)).astype(np.float32)
    n2 = 12
    out, bins = OL.oracle_binning(pv, n2)
    X = _normalised(pv)
    P = _normalised(np.hstack([pv, out]))[:, pv.shape[1]:]  # the picks in the same normalisation
    idx = [int(np.argmin(((X - P[:, [j]]) ** 2).sum(0))) for j in range(n2)]
    assert len(set(idx)) == n2 and np.allclose(X[:, idx], P, rtol=1e-6)
    mean = X.mean(axis=1, keepdims=True)
    d0 = ((X - mean) ** 2).sum(0)
    assert d0[idx[0]] >= d0.max() * (1 - 1e-5)
    d1 = ((X - X[:, [idx[0]]]) ** 2).sum(0)
    assert d1[idx[1]] >= d1.max() * (1 - 1e-5)
    D = np.array([((X - X[:, [i]]) ** 2).sum(0) for i in idx])
    assert np.all(D[bins - 1, np.arange(X.shape[1])] <= D.min(axis=0) * (1 + 1e-5) + 1e-12)


def test_oracle_small_known_answers():
    """Points 0, 1, 2, 10 on the mua axis (mus, g constant; range 10): the
    mean is 3.25, so the first pick is 10, then 0 (farthest from 10), then 2
    (nearest-pick distance 4/100 against 1/100 for set 1).  Labels: 10 -> 1,
    0 -> 2, 1 -> 2 (equally near 0 and 2; a label only moves on a strictly
    smaller distance), 2 -> 3."""
    pv = np.array([[0, 1, 2, 10], [5, 5, 5, 5], [0.9, 0.9, 0.9, 0.9]], np.float32)
    out, bins = OL.oracle_binning(pv, 3)
    np.testing.assert_array_equal(out[0], [10, 0, 2])
    np.testing.assert_array_equal(out[1:], np.array([[5, 5, 5], [0.9, 0.9, 0.9]], np.float32))
    np.testing.assert_array_equal(bins, [2, 2, 3, 1])
    # one bin: every label 1, the pick is the set farthest from the mean
    out1, bins1 = OL.oracle_binning(pv, 1)
    assert out1[0, 0] == 10 and np.all(bins1 == 1)
    # nothing varies: every pick is the first set, every label 1
    same = np.tile([[1.5], [7.0], [0.8]], (1, 9)).astype(np.float32)
    o, b = OL.oracle_binning(same, 4)
    assert np.all(o == same[:, :1]) and np.all(b == 1)


def test_oracle_rejects_empty():
    with pytest.raises(RuntimeError):
        OL.oracle_binning(np.zeros((3, 4), np.float32), 0)
