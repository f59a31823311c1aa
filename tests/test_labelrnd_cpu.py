"""Label randomisation (rndlabeler, src/PARAPROBE_SolverHdl.cpp:2809-3000) on
the host: the generator, the shuffle and the settings key, without a GPU."""
import numpy as np

from paraprobe_amd import horderdist as H


def _mt19937_py(seed, discard, count):
    """MT19937 as published by Matsumoto and Nishimura (init_genrand seeding,
    twist, tempering), written out independently of the C++ library."""
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    idx = 624
    out = []
    for _ in range(discard + count):
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y)
    return out[discard:]


def test_mt19937_standard_known_answer():
    # the C++ standard fixes the 10000th output of a default-seeded mt19937
    assert int(H.mt19937(5489, 9999, 1)[0]) == 4123659995


def test_mt19937_reference_seed_and_discard():
    got = [int(v) for v in H.mt19937(H.RND_SEED, H.RND_DISCARD, 4)]
    assert got == _mt19937_py(4294967295, 700000, 4)
    assert got != [int(v) for v in H.mt19937(H.RND_SEED, 0, 4)]


def test_shuffle_is_a_seeded_permutation():
    rng = np.random.default_rng(3)
    ty = rng.choice([1, 2, 3], size=5000, p=[0.7, 0.2, 0.1]).astype(np.uint8)
    a = H.shuffle_types(ty)
    assert np.array_equal(np.sort(a), np.sort(ty))     # counts per type kept
    assert not np.array_equal(a, ty)
    assert np.array_equal(a, H.shuffle_types(ty))      # deterministic
    assert not np.array_equal(a, H.shuffle_types(ty, discard=0))
    assert not np.array_equal(a, H.shuffle_types(ty, seed=1))
    # a pure permutation of positions: shuffling distinct ids and indexing
    # with them gives the same result as shuffling the types
    ids = np.arange(256, dtype=np.uint8)
    perm = H.shuffle_types(ids)
    assert np.array_equal(np.sort(perm), ids)
    small = ty[:256]
    assert np.array_equal(H.shuffle_types(small), small[perm])


def test_settings_key():
    base = "<Parameters><AnalysisSpatDistrType>1</AnalysisSpatDistrType>{}</Parameters>"
    key = "<SpatStatAdditionalLabelRandomization>{}</SpatStatAdditionalLabelRandomization>"
    assert H.label_randomization(base.format(key.format(1))) == 1
    assert H.label_randomization(base.format(key.format(0))) == 0
    assert H.label_randomization(base.format(key.format(2))) == 0
    assert H.label_randomization(base.format("")) == 0
