"""Writes the committed fixtures of tests/golden/ (run from the repo root).

- generator_fingerprints.json: digests of fixed-seed workload fields, so the seeded
  generators of ghcm_amd/csrc/field_gen.cpp cannot drift silently.
- ring256_fmm64.npz: the 256x256 seven-ring workload (seed 1, walls at speed 1e-3,
  centre source) and its FP64 FMM solution from the oracle, stored as float32.
- hand_kats.json: small grids whose solutions are derived by hand from the closed forms
  (1-D strips: prefix sums; constant cost from a corner: k*c on the axes,
  c (1 + sqrt(2)/2) on the first diagonal cell).
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import ghcm_amd as g  # noqa: E402
import oracle  # noqa: E402


def fingerprints():
    out = {}
    specs = [("ring256_s1", dict(kind="ring", w=256, h=256, seed=1, wall_speed=1e-3, scale=1.0)),
             ("ring200x150_s9", dict(kind="ring", w=200, h=150, seed=9, wall_speed=1e-4, scale=0.5)),
             ("smooth96_s2", dict(kind="smooth", w=96, h=80, seed=2, sigma=4.0)),
             ("points_4096_s3", dict(kind="points", w=4096, h=4096, seed=3, n=8))]
    for name, s in specs:
        if s["kind"] == "ring":
            f = g.gen_ring_speed(s["w"], s["h"], s["seed"], s["wall_speed"], s["scale"])
            s["sha256"] = hashlib.sha256(f.tobytes()).hexdigest()
        elif s["kind"] == "smooth":
            f = g.gen_smooth_speed(s["w"], s["h"], s["seed"], s["sigma"])
            s["sha256"] = hashlib.sha256(np.round(np.log10(f), 9).tobytes()).hexdigest()
        else:
            s["points"] = g.gen_points(s["w"], s["h"], s["seed"], s["n"]).tolist()
        out[name] = s
    return out


def main():
    with open(os.path.join(HERE, "generator_fingerprints.json"), "w") as f:
        json.dump(fingerprints(), f, indent=1)
    speed = g.gen_ring_speed(256, 256, 1, 1e-3, 1.0)
    t = oracle.fmm_f64(1.0 / speed, [(128, 128)])
    np.savez_compressed(os.path.join(HERE, "ring256_fmm64.npz"),
                        speed=speed.astype(np.float32), time=t.astype(np.float32),
                        seeds=np.array([[128, 128]], dtype=np.uint32))
    kats = {
        "strip": {"cost": [[0.5, 1.0, 2.0, 4.0, 8.0]], "seeds": [[0, 0]],
                  "time": [[0.0, 1.0, 3.0, 7.0, 15.0]]},
        "corner_c2": {"cost": [[2.0] * 3] * 3, "seeds": [[0, 0]],
                      "partial": {"0,1": 2.0, "0,2": 4.0, "1,0": 2.0, "2,0": 4.0,
                                  "1,1": 2.0 * (1 + math.sqrt(2) / 2)}},
        "two_seeds_strip": {"cost": [[1.0] * 7], "seeds": [[0, 0], [6, 0]],
                            "time": [[0.0, 1.0, 2.0, 3.0, 2.0, 1.0, 0.0]]},
        "walled": {"cost": [[1.0, "inf", 1.0]], "seeds": [[0, 0]],
                   "time": [[0.0, "inf", "inf"]]},
    }
    with open(os.path.join(HERE, "hand_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)


if __name__ == "__main__":
    main()
