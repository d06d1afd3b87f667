"""Writes the known-answer fixtures of this directory.

The reference ships no test vectors for its dense path, so these fixtures
come from a second, independent implementation written here in pure Python
(arbitrary-precision integers, brute-force loops) directly from the
reference's definitions:

  rng_kat.json     pixel_rng / next_u64 / uniform of mvs/src/stereo.rs:47-72:
                   seeds and the first 12 draws (as integers k with
                   uniform = k / 2^24) for a grid of (x, y, salt).
  census_kat.json  census_map of mvs/src/lib.rs:119-142 on small seeded
                   images (bit (dy+3)*7+(dx+3) set when the clamped
                   neighbour is darker than the centre).

Run:  python tests/golden/make_golden.py
"""
import json
import os
import random

M64 = (1 << 64) - 1
HERE = os.path.dirname(os.path.abspath(__file__))


def pixel_rng(x: int, y: int, salt: int) -> int:
    # `x as u64` of an i32 is its two's-complement sign extension
    s = ((x & M64) * 0x9E3779B97F4A7C15 ^ (y & M64) * 0xC2B2AE3D27D4EB4F ^ salt * 0x165667B19E3779F9) & M64
    s ^= s >> 30
    s = (s * 0xBF58476D1CE4E5B9) & M64
    s ^= s >> 27
    s = (s * 0x94D049BB133111EB) & M64
    return s ^ (s >> 31)


def xorshift(s: int) -> int:
    s ^= (s << 13) & M64
    s ^= s >> 7
    s ^= (s << 17) & M64
    return s & M64


def rng_kat():
    rows = []
    pts = [(0, 0, 7), (1, 0, 7), (0, 1, 7), (-1, -1, 0), (1919, 1079, 22), (2147483647, -2147483648, 3),
           (5, 9, 2**63 + 11)]
    r = random.Random(20240601)
    for _ in range(40):
        pts.append((r.randint(-100, 4000), r.randint(-100, 3000), r.randint(0, 40)))
    for (x, y, salt) in pts:
        s0 = pixel_rng(x, y, salt)
        s = s0
        draws = []
        for _ in range(12):
            s = xorshift(s)
            draws.append(s >> 40)
        rows.append({"x": x, "y": y, "salt": salt, "seed": str(s0), "u24": draws})
    return rows


def census(img, w, h):
    out = []
    for y in range(h):
        for x in range(w):
            c = img[y * w + x]
            bits = 0
            for dy in range(-3, 4):
                for dx in range(-3, 4):
                    if dx == 0 and dy == 0:
                        continue
                    xx = min(max(x + dx, 0), w - 1)
                    yy = min(max(y + dy, 0), h - 1)
                    if img[yy * w + xx] < c:
                        bits |= 1 << ((dy + 3) * 7 + (dx + 3))
            out.append(str(bits))
    return out


def census_kat():
    r = random.Random(7)
    cases = []
    for (w, h) in [(1, 1), (3, 2), (9, 11), (13, 5)]:
        img = [r.randint(0, 255) for _ in range(w * h)]
        if w * h > 10:
            img[0:4] = [9, 9, 9, 9]
        cases.append({"w": w, "h": h, "gray": img, "census": census(img, w, h)})
    return cases


if __name__ == "__main__":
    with open(os.path.join(HERE, "rng_kat.json"), "w") as f:
        json.dump(rng_kat(), f)
    with open(os.path.join(HERE, "census_kat.json"), "w") as f:
        json.dump(census_kat(), f)
    print("wrote rng_kat.json, census_kat.json")
