"""Counter RNG: Philox4x32-10 (Salmon et al., SC'11) pinned by the
Random123 known-answer vectors and an independent numpy restatement."""
import numpy as np
import pytest

import oracle_ffi as O

# Random123 kat_vectors, philox4x32 with 10 rounds: ctr, key -> out
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2,
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
     (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(O.philox(ctr, key)) == out


def philox_numpy(ctr, key, rounds=10):
    """Vectorised restatement: ctr (n,4) uint64 holding u32, key (2,)."""
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = 0x9E3779B9, 0xBB67AE85
    mask = np.uint64(0xFFFFFFFF)
    c = ctr.astype(np.uint64)
    k0, k1 = key
    for r in range(rounds):
        if r:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c[:, 0]
        p1 = M1 * c[:, 2]
        c = np.stack([(p1 >> np.uint64(32)) ^ c[:, 1] ^ np.uint64(k0),
                      p1 & mask,
                      (p0 >> np.uint64(32)) ^ c[:, 3] ^ np.uint64(k1),
                      p0 & mask], axis=1)
    return c.astype(np.uint32)


def test_philox_matches_numpy_restatement():
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, size=(200, 4), dtype=np.uint64)
    key = tuple(int(x) for x in rng.integers(0, 2**32, size=2))
    ref = philox_numpy(ctr, key)
    for i in range(len(ctr)):
        assert O.philox(tuple(int(x) for x in ctr[i]), key) == list(ref[i])


def test_uniform_mapping_and_counter_layout():
    seed = 0x0123456789ABCDEF
    pixel, sample, bounce, block = 77, 5, 3, 2
    u = O.uniform4(seed, pixel, sample, bounce, block)
    raw = O.philox((pixel, sample, (bounce << 20) | block, 0),
                   (seed & 0xFFFFFFFF, seed >> 32))
    expect = np.array([(x >> 8) * 2.0**-24 for x in raw], np.float32)
    assert np.array_equal(u, expect)
    assert np.all((u >= 0) & (u < 1))


def test_uniform_moments():
    us = np.array([O.uniform4(9, p, 0, 0, 1) for p in range(4000)]).ravel()
    assert abs(us.mean() - 0.5) < 0.01
    assert abs(us.var() - 1 / 12) < 0.003
