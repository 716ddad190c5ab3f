"""Derives the constant tables of FDLIBM 5.3's argument reduction (e_rem_pio2.c / k_rem_pio2.c) from pi itself.

pi is computed with Machin's formula in exact integer arithmetic, so the tables depend on no library and no copied
source. ``tests/test_oracle_fdlibm.py`` runs this to check the tables compiled into the oracle.
"""
import struct


def pi_fixed(bits: int) -> int:
    """floor(pi * 2**bits), from pi = 16 atan(1/5) - 4 atan(1/239) with 64 guard bits."""
    g = bits + 64
    one = 1 << g

    def atan_inv(k):
        total, term, n, sign = 0, one // k, 1, 1
        k2 = k * k
        while term:
            total += sign * (term // n)
            term //= k2
            n += 2
            sign = -sign
        return total

    return (16 * atan_inv(5) - 4 * atan_inv(239)) >> 64


def two_over_pi_chunks(count: int = 66):
    """The first ``count`` 24-bit chunks of the binary fraction of 2/pi (0.A2F983... in hex)."""
    bits = 24 * count + 64
    p = pi_fixed(bits + 8)
    frac = (2 << (2 * bits + 8)) // p          # 2/pi * 2**bits
    return [(frac >> (bits - 24 * (i + 1))) & 0xFFFFFF for i in range(count)]


def pio2_chunks(count: int = 8):
    """pi/2 split into doubles of 24 significant bits each (k_rem_pio2.c's PIo2[])."""
    p = pi_fixed(24 * count + 64)              # pi * 2**(24 count + 64)
    half = p >> 1                              # pi/2 at the same scale
    scale = 24 * count + 64
    out, rem = [], half
    exp = 1                                    # pi/2 = 1.57..., chunk 0 covers bits 2^0 .. 2^-23
    for i in range(count):
        shift = scale - (24 * i) + 1 - 24      # keep 24 bits starting at 2^(1 - 24 i - 1)
        chunk = rem >> shift
        rem -= chunk << shift
        out.append(chunk * 2.0 ** (shift - scale))
    return out


def npio2_hw(count: int = 32):
    """High 32 bits of the double nearest n*pi/2, n = 1..count (e_rem_pio2.c's npio2_hw[])."""
    p = pi_fixed(200)
    out = []
    for n in range(1, count + 1):
        v = n * p / 2 / 2 ** 200  # exact enough: float() rounds the big rational once
        num = n * p
        den = 2 * 2 ** 200
        d = num / den
        out.append(struct.unpack("<Q", struct.pack("<d", d))[0] >> 32)
    return out


if __name__ == "__main__":
    print("two_over_pi:", ", ".join(f"0x{c:06X}" for c in two_over_pi_chunks()))
    print("PIo2:", ", ".join(repr(v) for v in pio2_chunks()))
    print("PIo2 hex:", ", ".join(f"{struct.unpack('<Q', struct.pack('<d', v))[0]:016X}" for v in pio2_chunks()))
    print("npio2_hw:", ", ".join(f"0x{h:08X}" for h in npio2_hw()))


def _round_to_double(num: int, scale: int) -> float:
    """num / 2**scale rounded once to the nearest double."""
    from fractions import Fraction
    return float(Fraction(num, 1 << scale))


def medium_constants():
    """invpio2 and the 33+53-bit splits of pi/2 used by e_rem_pio2.c's medium-size path."""
    scale = 400
    half = pi_fixed(scale) >> 1                                  # pi/2 * 2**scale
    top = half.bit_length()                                      # bit length of the integer part + fraction
    out = {}
    rem = half
    for name in ("pio2_1", "pio2_2", "pio2_3"):
        keep = rem.bit_length() - 33                             # truncate to 33 significant bits
        head = (rem >> keep) << keep
        out[name] = _round_to_double(head, scale)
        rem -= head
        out[name + "t"] = _round_to_double(rem, scale)
    p = pi_fixed(scale)
    out["invpio2"] = _round_to_double((2 << (2 * scale)) // p, scale)
    return out


def bits(v: float) -> str:
    return f"{struct.unpack('<Q', struct.pack('<d', v))[0]:016X}"


if __name__ == "__main__":
    for k, v in medium_constants().items():
        print(k, bits(v), repr(v))
