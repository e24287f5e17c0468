"""Pins the oracle's counter-based RNG (Philox4x32-10, Salmon et al. SC'11)
against the published known-answer vectors of the Random123 distribution,
and checks basic uniformity of the walk streams."""
import numpy as np

import oracle_ffi as O

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key, expected).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        assert tuple(O.philox(ctr, key)) == want


def test_philox_outputs_uniform_bytes():
    words = []
    for i in range(4096):
        words.extend(O.philox((i, 0, 7, 0), (12345, 0)))
    b = np.frombuffer(np.array(words, dtype=np.uint32).tobytes(), dtype=np.uint8)
    counts = np.bincount(b, minlength=256)
    expected = len(b) / 256
    chi2 = float(((counts - expected) ** 2 / expected).sum())
    # 255 dof: mean 255, sd ~22.6; 400 is > 6 sd.
    assert chi2 < 400, chi2


def test_philox_streams_differ_by_walker_and_stream():
    a = O.philox((0, 0, 1, 0), (1, 0))
    b = O.philox((0, 0, 2, 0), (1, 0))
    c = O.philox((0, 0, 1, 1), (1, 0))
    assert a != b and a != c and b != c
