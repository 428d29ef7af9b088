"""Writes tests/golden/rng_first16.json: the first 16 raw 64-bit outputs and
uniform doubles of the input generator (xoshiro256** seeded by SplitMix64,
stream k of seed s starts SplitMix64 at s + k * 0xd1342543de82ef95), for the
seeds the benchmark configurations use.  Computed by this independent
pure-Python implementation of the published algorithms; the SplitMix64
anchor (state 0 -> 0xe220a8397b1dcdaf) and the xoshiro256** anchor (state
{1,2,3,4} -> 11520, 0) are checked first.
"""
import json
import os

M = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return x, z ^ (z >> 31)


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M


def xoshiro(s):
    r = (rotl((s[1] * 5) & M, 7) * 9) & M
    t = (s[1] << 17) & M
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = rotl(s[3], 45)
    return r


def stream(seed, k, n):
    x = (seed + k * 0xD1342543DE82EF95) & M
    s = []
    for _ in range(4):
        x, z = splitmix64(x)
        s.append(z)
    raw = [xoshiro(s) for _ in range(n)]
    return raw, [(r >> 11) * 2.0 ** -53 for r in raw]


def main():
    assert splitmix64(0)[1] == 0xE220A8397B1DCDAF
    s = [1, 2, 3, 4]
    assert xoshiro(s) == 11520 and xoshiro(s) == 0
    out = {}
    for seed in (0, 1, 2, 3, 12345):
        for k in (0, 1):
            raw, u = stream(seed, k, 16)
            out["%d:%d" % (seed, k)] = {"raw": [str(r) for r in raw], "u": u}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_first16.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
