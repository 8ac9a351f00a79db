"""Deterministic test sequences (numpy RNG): random ACGT, IUPAC/N runs,
lower case, low-complexity repeats, edge lengths."""
import numpy as np

ACGT = np.frombuffer(b"ACGT", np.uint8)


def random_seq(rng, n):
    return ACGT[rng.integers(0, 4, n)].tobytes().decode()


def mutate(rng, s, rate):
    a = np.frombuffer(s.encode(), np.uint8).copy()
    m = rng.random(len(a)) < rate
    a[m] = ACGT[rng.integers(0, 4, int(m.sum()))]
    return a.tobytes().decode()


def with_noise(rng, s):
    """Insert an N run, a few IUPAC codes and a lower-case stretch."""
    a = list(s)
    if len(a) > 40:
        p = int(rng.integers(0, len(a) - 20))
        for i in range(p, p + int(rng.integers(1, 15))):
            a[i] = "N"
        for _ in range(3):
            a[int(rng.integers(0, len(a)))] = "RYSWKM"[int(rng.integers(0, 6))]
        q = int(rng.integers(0, len(a) - 10))
        for i in range(q, q + 5):
            a[i] = a[i].lower()
    return "".join(a)


def mixed_set(seed, n, lmin, lmax):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        s = random_seq(rng, L)
        if i % 4 == 1:
            s = with_noise(rng, s)
        if i % 7 == 3:
            unit = random_seq(rng, int(rng.integers(1, 4)))
            s = (unit * (L // len(unit) + 1))[:L]
        out.append(s)
    return out
