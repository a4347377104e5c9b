"""Writes tests/golden/rng.json: WangHash and xorshift32 vectors computed with
plain Python integers (an independent statement of
lib/CUDA/shared_kernel_code/tools_shared.h:60-62 and lib/platform/system.cpp:46-52),
used to pin the oracle's and the kernels' RNG."""
import json
import os

M = 0xFFFFFFFF


def wang(s):
    s = (s ^ 61) ^ (s >> 16)
    s = (s * 9) & M
    s = s ^ (s >> 4)
    s = (s * 0x27D4EB2D) & M
    s = s ^ (s >> 15)
    return s


def xor32_seq(seed, n):
    out = []
    for _ in range(n):
        seed ^= (seed << 13) & M
        seed ^= seed >> 17
        seed ^= (seed << 5) & M
        out.append(seed)
    return out


def main():
    wang_in = [0, 1, 2, 61, 1234, 91771, 7907, 0x12345678, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF] + [i * 2654435761 & M for i in range(1, 33)]
    data = {
        "wang_hash": {"in": wang_in, "out": [wang(x) for x in wang_in]},
        "xor32": {str(s): xor32_seq(s, 64) for s in (1, 1234, 0x12345678)},
        # RandomFloat = xorshift * 2.3283064365387e-10f, first floats of seed WangHash(0 + R0), R0 = first xorshift of 1234
        "note": "xor32[seed] lists the first 64 outputs of RandomUInt(seed) starting from seed",
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
