"""Derives facts about the input files the reference ships (input/groups.csv
and input/particles.csv, written by its Blender add-on blender/seed_export.py
from a mesh that is not in the repository) and writes them to
tests/golden/seed_addon_facts.json: the files' SHA-256 and sizes, the row
count, the group's settings, and per Blender axis the float32 lattice origin
that reproduces every coordinate by the add-on's double-accumulated loop,
the spacing, and the range of lattice indices used.  No row of either file
is stored.  Run where the reference is checked out."""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as orc  # noqa: E402

SRC = "/root/reference/input"
S = 0.03


def points(path):
    rows = [l.split(",") for l in open(path).read().splitlines()[1:] if l]
    solver = np.array([[float(v) for v in r[1:]] for r in rows])
    return np.stack([solver[:, 0], -solver[:, 2], solver[:, 1]], axis=1)  # Blender's frame


def f32_near(v, k):
    out, a, b = [np.float32(v)], np.float32(v), np.float32(v)
    for _ in range(k):
        a, b = np.nextafter(a, np.float32(-np.inf)), np.nextafter(b, np.float32(np.inf))
        out += [a, b]
    return out


def axis_origin(vals):
    """The float32 lo whose lattice lo + s/2 + k s (accumulated in double,
    stored as float32) contains every value; (lo, index of min, index of max)."""
    s32 = float(np.float32(S))
    found = []
    for k0 in range(6):
        for lo in f32_near(min(vals) - s32 * 0.5 - k0 * s32, 200):
            seq = orc.seed_axis(float(lo), float(lo) + (len(vals) + k0 + 1) * s32, S).tolist()
            if vals <= set(seq):
                found.append((float(lo), seq.index(min(vals)), seq.index(max(vals))))
    assert len({f[0] for f in found}) == 1, found
    return found[0]


def facts(src=SRC):
    out = {"spacing": S}
    for name in ("groups.csv", "particles.csv"):
        data = open(os.path.join(src, name), "rb").read()
        out[name] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
    pts = points(os.path.join(src, "particles.csv"))
    out["rows"] = int(len(pts))
    g = open(os.path.join(src, "groups.csv")).read().splitlines()[1].split(",")
    out["group"] = {"velocity_solver": [float(v) for v in g[1:4]], "mass": float(g[4])}
    out["axes"] = []
    for e in range(3):
        vals = set(np.unique(pts[:, e]).tolist())
        lo, kmin, kmax = axis_origin(vals)
        out["axes"].append({"origin_f32": lo, "k_min": kmin, "k_max": kmax, "distinct": len(vals)})
    return out


if __name__ == "__main__":
    dst = os.path.join(HERE, "seed_addon_facts.json")
    with open(dst, "w") as f:
        json.dump(facts(), f, indent=1)
    print("wrote", dst)
