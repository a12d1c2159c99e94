"""Writes tests/golden/rng_vectors.json from independent pure-Python models
(big-int PCG32 per O'Neill's pcg32_srandom_r / pcg32_random_r, and TEA on
Python ints with explicit 32-bit wrap and arithmetic shift).  The oracle is
not used here.  The first entry is the published pcg32 demo sequence for
(initstate 42, initseq 54)."""
import json
import os

M64, M32 = 2 ** 64, 2 ** 32


def pcg32(initstate, initseq, n, index=0):
    state, inc = 0, (((initseq + index) % M64) << 1 | 1) % M64

    def step():
        nonlocal state
        old = state
        state = (old * 0x5851F42D4C957F2D + inc) % M64
        xs = (((old >> 18) ^ old) >> 27) % M32
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) % M32

    step()
    state = (state + initstate) % M64
    step()
    return [step() for _ in range(n)]


def tea(v0, v1, rounds=4):
    def s32(x):
        x %= M32
        return x - M32 if x >= 2 ** 31 else x

    v0, v1, s = v0 % M32, v1 % M32, 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) % M32
        v0 = (v0 + ((((v1 << 4) + 0xA341316C) % M32) ^ ((v1 + s) % M32) ^ (((s32(v1) >> 5) + 0xC8013EA4) % M32))) % M32
        v1 = (v1 + ((((v0 << 4) + 0xAD90777D) % M32) ^ ((v0 + s) % M32) ^ (((s32(v0) >> 5) + 0x7E95761E) % M32))) % M32
    return s32(v0), s32(v1)


def main():
    cases = [(42, 54, 0), (0x853C49E6748FEA9B, 0xDA3E39CB94B95BDB, 0), (7, 2 ** 64 - 3, 5), (123456789, 1, 2 ** 40)]
    out = {
        "pcg32_demo_42_54": ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b", "0xcbed606e"],
        "pcg32": [dict(initstate=a, initseq=b, index=k, out=pcg32(a, b, 8, k)) for a, b, k in cases],
        "tea": [dict(v0=a, v1=b, out=list(tea(a, b))) for a, b in
                [(0, 0), (0, 1), (12345, -7), (-1, 2 ** 31 - 1), (2 ** 31 - 1, -2 ** 31), (99, 1000000)]],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
