#!/usr/bin/env python3
"""Writes rng_vectors.json: known-answer vectors of the per-pixel generator.

The kernel's generator is java.util.Random's 48-bit LCG
(AbstractKernel.java:75-77, 438-440, 1101-1110):
    s' = (s * 0x5DEECE66D + 0xB) mod 2^48,  nextFloat = (s' >> 24) / 2^24.
The vectors are computed here with exact Python integer arithmetic from that
published definition (java.util.Random, JDK API documentation), independently
of the oracle and of the GPU code.
"""

import json
import os

MULT = 0x5DEECE66D
ADD = 0xB
MASK = (1 << 48) - 1


def draws(seed, count):
    s = seed & ((1 << 64) - 1)
    out = []
    for _ in range(count):
        s = (s * MULT + ADD) & MASK
        out.append((s >> 24) / float(1 << 24))
    return out, s


def main():
    vectors = []
    for seed in (0, 42, -7364102858154096381, 0x123456789ABC, (1 << 63) - 1):
        vals, final = draws(seed, 1000)
        vectors.append({"seed": seed, "draws": vals, "final_seed": final})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump({"generator": "java.util.Random LCG, raw seeds (no scrambling)", "vectors": vectors}, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
