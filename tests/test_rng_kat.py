"""Philox4x32-10 known-answer vectors (Random123 kat_vectors) for the CPU
restatement that checks the device generator."""
import numpy as np

from rng_ref import normals, philox4x32_10


def test_philox_kat_zero():
    assert philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)


def test_philox_kat_ones():
    F = 0xFFFFFFFF
    assert philox4x32_10((F, F, F, F), (F, F)) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)


def test_philox_kat_pi():
    ctr = (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344)
    key = (0xA4093822, 0x299F31D0)
    assert philox4x32_10(ctr, key) == (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_box_muller_moments():
    z = np.concatenate([normals(7, p, 3, 12) for p in range(2000)])
    assert abs(z.mean()) < 0.02
    assert abs(z.std() - 1.0) < 0.02
