"""Writes tests/golden/rng_streams.json: xoshiro256++ streams of rng.zig
restated independently with Python integers (no shared code with the C++
oracle or the HIP kernels), for the oracle and device tests to compare with.

rng.zig:13-38 and 82-88: init(seed) fills the state with four SplitMix64
outputs; forStream(seed, s) = init(splitMix64(seed ^ (s * 0x9E3779B97F4A7C15)));
next() is xoshiro256++ (rotl 23, shift 17, rotl 45).
Run: python tests/golden/make_rng_fixture.py
"""
import json
import os

M = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix(state):
    state = (state + GOLDEN) & M
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return state, z ^ (z >> 31)


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M


def init(seed):
    s, out = seed, []
    for _ in range(4):
        s, v = splitmix(s)
        out.append(v)
    return out


def for_stream(seed, stream):
    _, v = splitmix(seed ^ ((stream * GOLDEN) & M))
    return init(v)


def next_(s):
    result = (rotl((s[0] + s[3]) & M, 23) + s[0]) & M
    t = (s[1] << 17) & M
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = rotl(s[3], 45)
    return result


def main():
    seeds = [0, 1, 42, M]
    streams = [0, 1, 159999, 1 << 63]
    doc = {"splitmix64_seed0": [], "init": {}, "for_stream": {}}
    st = 0
    for _ in range(3):
        st, v = splitmix(st)
        doc["splitmix64_seed0"].append(v)
    for seed in seeds:
        s = init(seed)
        doc["init"][str(seed)] = [next_(s) for _ in range(8)]
        for stream in streams:
            s = for_stream(seed, stream)
            doc["for_stream"][f"{seed}:{stream}"] = [next_(s) for _ in range(8)]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_streams.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main()
