"""Input/output vectors of the reference's media binning MEX
(+MCmatlab/src/getDownsampledParamVals.c), from the reference compiled from
source against the mx shim (oracle/_ref/libref_binning.so, built by
`make -C oracle ref`), into tests/golden/binning_vectors.npz.

Data only: for each case the 3 x N1 single input, N2, and the reference's
3 x N2 single output and uint8 labels.  The cases cover distinct sets, many
ties (small integer grids), rows that do not vary, nothing varying, one
input set, more bins than distinct sets and labels past 255 (which wrap).
Run in a container that has the reference:
    python tests/golden/make_binning_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as OL  # noqa: E402

OUT = os.path.join(HERE, "binning_vectors.npz")


def cases():
    rng = np.random.default_rng(20)
    yield "distinct", rng.random((3, 300)), 24
    yield "ties", rng.integers(0, 4, (3, 500)), 30
    g = rng.random((3, 400))
    g[2] = 0.9
    yield "g_constant", g, 17
    yield "nothing_varies", np.tile([[2.0], [150.0], [0.9]], (1, 50)), 5
    yield "one_set", np.array([[3.0], [100.0], [0.8]]), 4
    yield "more_bins_than_sets", rng.integers(0, 3, (3, 60)) * 0.37, 40
    yield "wrapping_labels", rng.random((3, 700)), 300
    mua = 10 + 5 * rng.random(2000)
    yield "mua_only", np.vstack([mua, np.full(2000, 200.0), np.full(2000, 0.9)]), 64


def main(out=OUT):
    data = {}
    for k, (name, pv, n2) in enumerate(cases()):
        pv = np.asarray(pv, dtype=np.float32)
        ds, bins = OL.ref_binning(pv, n2)
        data[f"{k}_{name}_pv"] = pv
        data[f"{k}_{name}_n2"] = np.array(n2)
        data[f"{k}_{name}_out"] = ds.astype(np.float32)
        data[f"{k}_{name}_bin"] = bins.astype(np.uint8)
    np.savez_compressed(out, **data)
    print("wrote", out, len(data) // 4, "cases")


if __name__ == "__main__":
    main()
