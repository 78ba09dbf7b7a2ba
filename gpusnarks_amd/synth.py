"""Seeded synthetic NTT inputs (SplitMix64 + rejection sampling).

Every test and benchmark draws its field elements from here, so the GPU path,
the oracle and the golden fixtures all see the same operands for a seed.
Elements are uniform residues in [0, p), returned as little-endian u32 limbs;
they are used directly as Montgomery-form representations (a uniform residue
is a uniform Montgomery representation).
"""
import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64_stream(seed, count):
    """`count` successive SplitMix64 outputs (numpy uint64) for `seed`."""
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_residues(seed, n, modulus, nlimbs):
    """n residues uniform in [0, modulus) as an (n, nlimbs) uint32 array.

    Each candidate takes ceil(bits/64) SplitMix64 words, masked to the
    modulus bit length; candidates >= modulus are rejected.
    """
    bits = modulus.bit_length()
    words = (bits + 63) // 64
    out = np.zeros((n, nlimbs), dtype=np.uint32)
    filled = 0
    stream_pos = 0
    state_seed = seed
    mod_limbs = [(modulus >> (32 * i)) & 0xFFFFFFFF for i in range(nlimbs)]
    top_mask = (1 << (bits - 32 * (nlimbs - 1))) - 1 if nlimbs > 1 else (1 << bits) - 1
    while filled < n:
        want = int((n - filled) * 1.3) + 16
        raw = splitmix64_stream(state_seed + stream_pos, want * words)
        stream_pos += want * words + 1
        cand = raw.reshape(want, words)
        limbs = np.zeros((want, 2 * words), dtype=np.uint32)
        limbs[:, 0::2] = (cand & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        limbs[:, 1::2] = (cand >> np.uint64(32)).astype(np.uint32)
        limbs = limbs[:, :nlimbs].copy()
        limbs[:, nlimbs - 1] &= np.uint32(top_mask)
        # keep candidates < modulus (lexicographic compare from the top limb)
        lt = np.zeros(want, dtype=bool)
        eq = np.ones(want, dtype=bool)
        for i in range(nlimbs - 1, -1, -1):
            m = np.uint32(mod_limbs[i])
            lt |= eq & (limbs[:, i] < m)
            eq &= limbs[:, i] == m
        good = limbs[lt]
        take = min(len(good), n - filled)
        out[filled:filled + take] = good[:take]
        filled += take
    return out


def to_int(limbs_row):
    return sum(int(v) << (32 * i) for i, v in enumerate(limbs_row))


def from_int(x, nlimbs):
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(nlimbs)], dtype=np.uint32)


# Field ids of the C ABI (include/gpusnarks.h); repeated here so this module
# imports without the library.
_FP32, _MNT6 = 0, 2


def ragged_workload(seed=0x5EED0003, count=256):
    """The batched ragged workload: `count` (field, log2 n, inverse) items,
    75% over MNT6-753 Fr and 25% over the 32-bit field, log2 n uniform over
    the integers [10, 18], forward or inverse 50/50, in shuffled order."""
    import random
    rng = random.Random(seed)
    items = []
    for i in range(count):
        field = _MNT6 if i < count * 3 // 4 else _FP32
        items.append((field, rng.randint(10, 18), rng.random() < 0.5))
    rng.shuffle(items)
    return items


RAGGED_INPUT_SEED = 0xBA7C4  # item k of the ragged workload draws seed RAGGED_INPUT_SEED + k


def serving_requests(seed=0x5EED0004, count=1000):
    """The serving mix: `count` (field, log2 n, inverse) requests, 70%
    MNT6-753 Fr and 30% 32-bit field, log2 n uniform over the integers
    [12, 20], forward or inverse 50/50."""
    import random
    rng = random.Random(seed)
    reqs = []
    for _ in range(count):
        field = _MNT6 if rng.random() < 0.7 else _FP32
        reqs.append((field, rng.randint(12, 20), rng.random() < 0.5))
    return reqs
