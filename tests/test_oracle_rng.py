"""Known-answer tests of the two random number generators.

dSFMT-19937 (the reference CPU path's RNG, MCmatlab.c:144-147, :241): the
oracle's restatement against the generator's own published check file
(dSFMT.19937.out.txt, shipped with the reference, committed as
golden/dsfmt19937_kat.json), which prints 15 decimals.

Philox4x32-10 (the GPU path's RNG): the oracle's and the device header's
implementation against the known-answer vectors published with Random123
(Salmon et al., SC'11): counter/key all zeros, all ones, and the pi digits.
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle_lib as OL

HERE = os.path.dirname(os.path.abspath(__file__))

PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0], [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def _kat():
    with open(os.path.join(HERE, "golden", "dsfmt19937_kat.json")) as f:
        return json.load(f)["sections"]


def _open_open(x12):
    """(0,1) variant of the generator: set the lowest mantissa bit of the
    [1,2) draw, subtract 1."""
    b = struct.unpack("<Q", struct.pack("<d", x12))[0] | 1
    return struct.unpack("<d", struct.pack("<Q", b))[0] - 1.0


@pytest.mark.parametrize("sec", _kat(), ids=lambda s: f"seed{s['seed']}_{s['interval']}")
def test_dsfmt_matches_published_check_file(sec):
    n = len(sec["values"])
    oc = OL.dsfmt_stream(sec["seed"], n)        # (0,1] = 2 - [1,2)
    c12 = 2.0 - oc
    want = np.array(sec["values"])
    got = {"[1, 2)": c12, "[0, 1)": c12 - 1.0, "(0, 1]": oc, "(0, 1)": np.array([_open_open(v) for v in c12])}[sec["interval"]]
    # the file prints %1.15f: agreement to half a unit in the 15th decimal
    np.testing.assert_allclose(got, want, rtol=0, atol=5.01e-16)


def test_dsfmt_long_stream_is_the_first_section():
    # the first section holds 1000 values, crossing two state refills (382 each)
    sec = _kat()[0]
    assert len(sec["values"]) == 1000 and sec["interval"] == "[1, 2)"
    got = 2.0 - OL.dsfmt_stream(sec["seed"], 1000)
    np.testing.assert_allclose(got, sec["values"], rtol=0, atol=5.01e-16)


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert OL.philox(ctr, key) == want


def test_philox_device_header_matches_oracle():
    """The GPU's Philox header, compiled for the host in the test harness,
    gives the same words as the oracle for arbitrary counters."""
    import ctypes as C

    path = os.path.join(HERE, "_build", "libmc_host.so")
    if not os.path.exists(path):
        pytest.skip("host harness not built")
    lib = C.CDLL(path)
    lib.mcm_test_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    rng = np.random.default_rng(5)
    for _ in range(50):
        ctr = [int(v) for v in rng.integers(0, 2 ** 32, 4)]
        key = [int(v) for v in rng.integers(0, 2 ** 32, 2)]
        o = (C.c_uint32 * 4)()
        lib.mcm_test_philox((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), o)
        assert list(o) == OL.philox(ctr, key)
