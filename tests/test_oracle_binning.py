"""The media-binning oracle (oracle/binning_oracle.c) against hand-worked
answers of the reference's algorithm (+MCmatlab/src/getDownsampledParamVals.c)
and brute-force properties of farthest-point sampling that a plausible
mistake (a wrong distance, a lost tie rule, labels off by one) breaks.

The reference's MEX source needs MATLAB's mex.h and mx libraries, so it is
not built here; the worked cases below are chosen so that every distance is
exact in float arithmetic (each varying row spans a power of two), and each
case's docstring line states the steps that fix its answer."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as OL


def _pv(mua, mus=None, g=None):
    n = len(mua)
    return np.array([mua, mus if mus is not None else [5.0] * n, g if g is not None else [0.9] * n], np.float32)


def worked_cases():
    """(name, paramVals 3 x N1, N2, downsampledParamVals 3 x N2, binidx)

    scan_rule: mua 1, 10, 7, 14, 0, 16 (range 16, mean 8).  Pick 1 = set 4
      (0; tied with 16 at distance 8 from the mean, the first maximum wins).
      Pick 2 = set 5 (16, farthest from 0).  Pass 1 (running index starts at
      0): nearest-pick squared distances 1, 36, 49, 4, 0, 0 -> pick 3 = set 2
      (7).  Pass 2: new distances 1, 9, 0, 4, 0, 0; the running index starts
      at set 2 and only moves to a set that exceeds the distance it is
      compared with: sets 0 and 1 are compared with set 2's old distance 49,
      set 3 with set 2's new distance 0 -> running index 3; 4 > d[0] = 1, so
      pick 4 = set 3 (14), although set 1 (10, distance 3) is farther.
      Labels: 1, 3, 3, 4, 1, 2.
    repeat_pick: mua 9, 11, 16, 0, 12 (range 16).  Picks 0 (set 3), 16
      (set 2), 9 (set 0: the running index stays at 0 and d[0] = 49 is the
      maximum), 12 (set 4); pass 3 starts at set 4 (old distance 9 > every
      new distance before it), set 4's new distance is 0 and nothing after
      it exceeds 0, and d[4] = 0 is not > d[0] = 0, so pick 5 = set 0 again
      (9), not set 1 (11).  Labels 3, 4, 2, 1, 4.
    two_rows: (mua, mus) = (0,0), (8,0), (0,8), (8,8), (4,4), g constant.
      Normalised: the corners of the unit square and its centre (mean).  The
      four corners are equally far from the mean: pick 1 = set 0; pick 2 =
      set 3 (distance 2); pass 1: distances 0, 1, 1, 0, 1/2 -> set 1; pass 2
      starts at set 1 (old distance 1): set 2 is compared with set 1's new
      distance 0 -> pick 4 = set 2.  Set 4 is equally near every pick and
      keeps label 1.  Labels 1, 3, 4, 2, 1.
    small: mua 0, 1, 2, 10 (range 10, mean 3.25): picks 10, 0, 2; set 1 is
      equally near 0 and 2 and keeps label 2 (a label moves only on a
      strictly smaller distance).  Labels 2, 2, 3, 1.
    one_bin: the same sets, one bin: the pick is the set farthest from the
      mean (10), every label 1.
    constant: nothing varies: every pick is the first set, every label 1.
    """
    f32 = lambda a: np.array(a, np.float32)
    return [
        ("scan_rule", _pv([1, 10, 7, 14, 0, 16]), 4, f32([[0, 16, 7, 14], [5] * 4, [0.9] * 4]), np.uint8([1, 3, 3, 4, 1, 2])),
        ("repeat_pick", _pv([9, 11, 16, 0, 12]), 5, f32([[0, 16, 9, 12, 9], [5] * 5, [0.9] * 5]), np.uint8([3, 4, 2, 1, 4])),
        ("two_rows", _pv([0, 8, 0, 8, 4], [0, 0, 8, 8, 4]), 4, f32([[0, 8, 8, 0], [0, 8, 0, 8], [0.9] * 4]),
         np.uint8([1, 3, 4, 2, 1])),
        ("small", _pv([0, 1, 2, 10]), 3, f32([[10, 0, 2], [5] * 3, [0.9] * 3]), np.uint8([2, 2, 3, 1])),
        ("one_bin", _pv([0, 1, 2, 10]), 1, f32([[10], [5], [0.9]]), np.uint8([1, 1, 1, 1])),
        ("constant", np.tile([[1.5], [7.0], [0.8]], (1, 9)).astype(np.float32), 4, np.tile([[1.5], [7.0], [0.8]], (1, 4)).astype(np.float32),
         np.ones(9, np.uint8)),
    ]


@pytest.mark.parametrize("case", worked_cases(), ids=lambda c: c[0])
def test_oracle_matches_worked_cases(case):
    _, pv, n2, out, bins = case
    o, b = OL.oracle_binning(pv, n2)
    assert np.array_equal(o.view(np.uint32), out.view(np.uint32)), o
    assert np.array_equal(b, bins), b


def test_scan_rule_differs_from_plain_farthest_point():
    """Plain farthest-point sampling (a fresh first-maximum scan per pass)
    would pick 10 fourth in the scan_rule case; the reference's running
    index picks 14.  A restatement that resets the running index fails."""
    name, pv, n2, out, bins = worked_cases()[0]
    x = pv[0].astype(np.float64)
    d = np.minimum.reduce([(x - v) ** 2 for v in (0, 16, 7)])
    assert x[int(np.argmax(d))] == 10 and out[0, 3] == 14


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
    farthest set; the worked cases pin those."""
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


def test_oracle_rejects_empty():
    with pytest.raises(RuntimeError):
        OL.oracle_binning(np.zeros((3, 4), np.float32), 0)
