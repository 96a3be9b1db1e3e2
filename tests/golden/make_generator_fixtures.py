"""Writes tests/golden/generator_fixtures.json: known entries of DPLASMA's
test-matrix generators, computed independently of the C oracle and of the
GPU kernels.

The reference ships no golden vectors for this path, so the expected values
come from the generator's definition (src/cores/random.h:19-41,
src/cores/core_zplgsy.c:72-150, src/cores/core_zplrnt.c:68-91, real case):
  X_0 = seed, X_{t+1} = 6364136223846793005 X_t + 1 (mod 2^64), computed here
  by stepping the recurrence (no jump-ahead), and
  entry = fl32(0.5 - fl32(X_t) * 2^-64), evaluated with exact rationals and
  round-to-nearest-even into binary32; the symmetric matrix adds the bump to
  the diagonal in binary64 and mirrors the lower triangle.
Run:  python tests/golden/make_generator_fixtures.py
"""
import json
import os
from fractions import Fraction

A_LCG = 6364136223846793005
C_LCG = 1
MASK = (1 << 64) - 1


def round_to_f32(fr):
    """Nearest binary32 value (ties to even) of a rational, as an exact Fraction."""
    if fr == 0:
        return Fraction(0)
    sign = -1 if fr < 0 else 1
    fr = abs(fr)
    # exponent e with 2^e <= fr < 2^(e+1)
    e = fr.numerator.bit_length() - fr.denominator.bit_length()
    if Fraction(2) ** e > fr:
        e -= 1
    while Fraction(2) ** (e + 1) <= fr:
        e += 1
    scaled = fr / (Fraction(2) ** (e - 23))       # in [2^23, 2^24)
    q, r = divmod(scaled.numerator, scaled.denominator)
    rem = Fraction(r, scaled.denominator)
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and q % 2 == 1):
        q += 1
    return sign * q * Fraction(2) ** (e - 23)


def entry(x):
    """0.5f - (float)x * RndF_Mul in binary32 (RndF_Mul == 2^-64)."""
    fx = round_to_f32(Fraction(x))
    s = round_to_f32(fx * Fraction(1, 1 << 64))    # exact scaling, kept for clarity
    return float(round_to_f32(Fraction(1, 2) - s))


def lcg_states(seed, count):
    out = [seed]
    x = seed
    for _ in range(count - 1):
        x = (A_LCG * x + C_LCG) & MASK
        out.append(x)
    return out


def plgsy_entry(states, N, i, j, bump):
    """A(i,j) of dplgsy(bump, N, seed) from the precomputed states X_0..X_t."""
    lo, hi = (i, j) if i >= j else (j, i)
    v = entry(states[lo + hi * N])
    if i == j:
        v = v + bump          # binary64 addition
    return v


def main():
    fixtures = {"description": __doc__.split("\n\n")[0], "cases": []}
    # small complete matrices (every entry), both generators
    for (N, seed) in [(8, 3872), (13, 7)]:
        st = lcg_states(seed, N * N)
        fixtures["cases"].append({
            "kind": "dplgsy_full", "N": N, "bump": float(N), "seed": seed,
            "values": [[plgsy_entry(st, N, i, j, float(N)).hex() for j in range(N)] for i in range(N)],
        })
    # selected entries of the configurations the testers use
    for (N, seed) in [(378, 3872), (4096, 3872)]:
        pos = [(0, 0), (1, 0), (N - 1, 0), (N - 1, N - 1), (N // 2, N // 3), (93, 92), (255, 256),
               (256, 255), (377, 200), (17, 17)]
        pos = [(i, j) for (i, j) in pos if i < N and j < N]
        tmax = max(max(i, j) + min(i, j) * N for (i, j) in pos) + 1
        st = lcg_states(seed, tmax)
        fixtures["cases"].append({
            "kind": "dplgsy_entries", "N": N, "bump": float(N), "seed": seed,
            "entries": [[i, j, plgsy_entry(st, N, i, j, float(N)).hex()] for (i, j) in pos],
        })
    # right-hand sides: dplrnt(M x NRHS, seed)
    for (M, nrhs, seed) in [(378, 2, 3873), (4096, 1, 3873)]:
        st = lcg_states(seed, M * nrhs)
        fixtures["cases"].append({
            "kind": "dplrnt", "M": M, "N": nrhs, "seed": seed,
            "values": [[entry(st[i + j * M]).hex() for j in range(nrhs)] for i in range(M)],
        })
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generator_fixtures.json")
    with open(out, "w") as f:
        json.dump(fixtures, f, indent=0)
    print("wrote", out)


if __name__ == "__main__":
    main()
