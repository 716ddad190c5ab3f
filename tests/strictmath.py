"""Argument generator and digest of the reference's StrictMath vector file, restated from its documented format
(tests/GridSim.Tests/FdlibmStrictMathVectorTests.cs:75-160: SplitMix64 arguments per class, SHA-256 over the
little-endian bits of each result, NaN canonicalised)."""
import hashlib
import math
import struct

import numpy as np

M64 = (1 << 64) - 1
CANON_NAN = 0x7FF8000000000000


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first ``count`` outputs of SplitMix64 from ``seed`` (vectorised: state_k = seed + k * golden)."""
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & M64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def draw(spec: str, seed: int, count: int) -> np.ndarray:
    kind, _, rest = spec.partition(":")
    if kind == "bits":
        return splitmix64(seed, count).view(np.float64)
    if kind == "sub":
        a = splitmix64(seed, count)
        v = (a & np.uint64(0x000FFFFFFFFFFFFF)).view(np.float64)
        return np.where(a.view(np.int64) < 0, -v, v)
    if kind == "r":
        lo, hi = (float(t) for t in rest.split(":"))
        u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u
    if kind == "lu":
        e1, e2 = (int(t) for t in rest.split(":"))
        z = splitmix64(seed, 2 * count).reshape(count, 2)
        a, b = z[:, 0], z[:, 1]
        m = 1.0 + (a >> np.uint64(12)).astype(np.float64) * (1.0 / 4503599627370496.0)
        e = e1 + ((b >> np.uint64(33)) % np.uint64(e2 - e1 + 1)).astype(np.int64)
        v = np.ldexp(m, e)
        return np.where((a & np.uint64(1)) != 0, -v, v)
    if kind == "nearpio2":
        kk = int(rest)
        z = splitmix64(seed, 2 * count).reshape(count, 2)
        n = ((z[:, 0] >> np.uint64(11)) % np.uint64(2 * kk + 1)).astype(np.int64) - kk
        x = n.astype(np.float64) * (math.pi / 2)
        j = ((z[:, 1] >> np.uint64(11)) % np.uint64(7)).astype(np.int64) - 3
        bits = x.view(np.int64).copy()
        out = x.copy()
        for i in range(count):
            v = x[i]
            for _ in range(abs(int(j[i]))):
                v = np.nextafter(v, math.inf if j[i] > 0 else -math.inf)
            out[i] = v
        del bits
        return out
    raise ValueError(spec)


def digest(values: np.ndarray) -> str:
    b = np.ascontiguousarray(values, dtype=np.float64).view(np.uint64).copy()
    b[np.isnan(values)] = np.uint64(CANON_NAN)
    return hashlib.sha256(b.astype("<u8").tobytes()).hexdigest()


def load(path):
    vec, dig = [], []
    for line in open(path, encoding="utf-8"):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "v":
            vec.append((t[1], int(t[2], 16), int(t[3], 16)))
        elif t[0] == "d":
            dig.append((t[1], t[2], int(t[3]), int(t[4]), t[5]))
    return vec, dig


def from_bits(u: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", u))[0]
