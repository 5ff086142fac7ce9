"""Philox4x32-10 known-answer vectors (Random123 distribution, kat_vectors)."""

from cs_afqmc_amd import rng


def test_philox_known_answers():
    assert rng.philox4x32_10((0, 0, 0, 0), (0, 0)) == (
        0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert rng.philox4x32_10((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert rng.philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
                             (0xA4093822, 0x299F31D0)) == (
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_split_and_uniform():
    k = rng.PRNGKey(5)
    k2, sub = rng.split(k)
    assert k2 != sub and k2[0] == sub[0] == 5
    u = [rng.uniform(rng.split((5, i))[1]) for i in range(200)]
    assert all(0.0 <= x < 1.0 for x in u)
    assert len(set(u)) == 200
