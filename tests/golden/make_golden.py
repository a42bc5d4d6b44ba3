"""Writes the known-answer fixtures of tests/golden/ from independent
restatements: integer arithmetic for the RNG streams and double-precision
closed forms for the Disney BSDF's textbook reductions.  Nothing here imports
the oracle or the library, so the fixtures pin both.

    python tests/golden/make_golden.py      (rewrites the .json files beside it)

rng_streams.json: for (pixel, sample, dimension) keys, the stream seed of
pathtracer.cpp:20-34 (hash(pixel + hash(sample + hash(dim + 0x9e3779b9))), 0
mapped to 1, rng.h:13) and the first 16 xorshift32 outputs of rng.h:15-29 with
their floats (x >> 8) * 2^-24.

bsdf_closed_forms.json: f * cos(theta_i) and pdf at fixed directions (normal
+z) for two material families whose Disney BSDF reduces to a textbook form
(disney_bsdf.h:95-137,148-175,200-234):
  lambert: specular = metallic = transmission = sheen = clearcoat =
    subsurface = 0: f*cos = base * NoL / pi, pdf = NoL / pi;
  metal: metallic = 1, base = 1, no clearcoat: F = 1,
    f*cos = D G / (4 NoV), pdf = D NoH / (4 VoH), with the GGX D and Smith G1
    at alpha = roughness.
"""
import json
import math
import os

M32 = 0xFFFFFFFF
HERE = os.path.dirname(os.path.abspath(__file__))


def hash_u32(x):
    x &= M32
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def xorshift(state, n):
    out = []
    for _ in range(n):
        state ^= (state << 13) & M32
        state ^= state >> 17
        state ^= (state << 5) & M32
        out.append(state)
    return out


def rng_streams():
    keys = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (7, 3, 1), (65535, 63, 9), (1048575, 1023, 255),
            (123456, 7, 0x10000 + 4), (3, 2, 0x10000)]
    rows = []
    for p, s, d in keys:
        seed = hash_u32(p + hash_u32(s + hash_u32(d + 0x9E3779B9)))
        seed = seed or 1
        u = xorshift(seed, 16)
        rows.append({"pixel": p, "sample": s, "dim": d, "seed": seed, "u32": u,
                     "f32": [(v >> 8) / 16777216.0 for v in u]})
    return rows


def normalize(v):
    n = math.sqrt(sum(x * x for x in v))
    return [x / n for x in v]


def dir_from(theta_deg, phi_deg):
    t, p = math.radians(theta_deg), math.radians(phi_deg)
    return [math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)]


def ggx_D(noh, a):
    a2 = a * a
    d = noh * noh * (a2 - 1) + 1
    return a2 / (math.pi * d * d)


def smith_G1(nov, a):
    a2 = a * a
    return 2 * nov / (nov + math.sqrt(a2 + (1 - a2) * nov * nov))


def bsdf_cases():
    pairs = [((10, 0), (40, 180)), ((35, 20), (35, 200)), ((60, 90), (15, 270)), ((5, 0), (80, 45)),
             ((45, 0), (45, 180)), ((70, 300), (20, 120))]
    rows = []
    for base in ((0.8, 0.5, 0.2), (0.1, 0.9, 0.4)):
        for (to, po), (ti, pi_) in pairs:
            wo, wi = dir_from(to, po), dir_from(ti, pi_)
            nol = wi[2]
            rows.append({"kind": "lambert", "base": list(base), "roughness": 1.0, "wo": wo, "wi": wi,
                         "f": [b * nol / math.pi for b in base], "pdf": nol / math.pi})
    for a in (0.15, 0.4, 0.8):
        for (to, po), (ti, pi_) in pairs:
            wo, wi = dir_from(to, po), dir_from(ti, pi_)
            h = normalize([wo[k] + wi[k] for k in range(3)])
            nov, nol, noh = wo[2], wi[2], h[2]
            voh = sum(wo[k] * h[k] for k in range(3))
            D = ggx_D(noh, a)
            G = smith_G1(nov, a) * smith_G1(nol, a)
            f = D * G / (4 * nov)
            rows.append({"kind": "metal", "base": [1.0, 1.0, 1.0], "roughness": a, "wo": wo, "wi": wi,
                         "f": [f, f, f], "pdf": D * noh / (4 * voh)})
    return rows


if __name__ == "__main__":
    with open(os.path.join(HERE, "rng_streams.json"), "w") as fh:
        json.dump(rng_streams(), fh, indent=0)
    with open(os.path.join(HERE, "bsdf_closed_forms.json"), "w") as fh:
        json.dump(bsdf_cases(), fh, indent=0)
