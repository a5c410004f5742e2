"""snowmpm.seed -- seed snow particles inside closed meshes and write the
solver's input files, without Blender.

Host side of the reference's Blender add-on (blender/seed_export.py).  Each
snow group is a closed triangle mesh (Wavefront OBJ) with the add-on's
per-object settings (SnowGroupSettings, seed_export.py:16-39: velocity,
particle mass, particle spacing).  The particles are seeded on the GPU
(``snowmpm.seed_mesh``, the add-on's seed_mesh / point_inside_mesh /
ray_inside), then written as ``groups.csv`` and ``particles.csv``
(export_groups / export_particles, seed_export.py:165-255), which
``snowmpm.import_particles`` and the reference's importer read.

What follows the add-on:
  * groups are ordered by object name (get_snow_groups, 148-163); the name is
    the OBJ file's first ``o`` statement, else the file's stem;
  * settings are float32 values, as Blender stores float properties (a mass
    of 0.05 is written as 0.05000000074505806);
  * coordinates convert from Blender's z-up frame to the solver's y-up frame
    as (x, z, -y) (blender_to_solver, 142-147), for positions and velocities,
    when the mesh is given in Blender's frame (``frame="blender"``, the
    default; ``frame="solver"`` takes the mesh as already y-up);
  * csv rows use Python's float repr and "\\n" line ends (csv.writer with
    lineterminator "\\n", 172, 219);
  * the output directory is created with its parents (309-312).
Polygons with more than three corners are split into a fan from their first
corner; Blender may split a quad along the other diagonal, which matters
only for points within the non-planarity of that quad.

Command line:
  python -m snowmpm.seed OUTDIR --group MESH.obj [velocity=VX,VY,VZ]
                                   [mass=M] [spacing=S] [--group ...]
                                   [--frame blender|solver] [--device N]
"""
import argparse
import csv
import sys
from pathlib import Path

import numpy as np

from . import seed_mesh

# SnowGroupSettings defaults (seed_export.py:21-38)
DEFAULT_VELOCITY = (0.0, 0.0, 0.0)
DEFAULT_MASS = 0.05
DEFAULT_SPACING = 0.05


def read_obj(path):
    """Vertices (n x 3 float64), triangles (k x 3 int32) and object name of a
    Wavefront OBJ file.  Faces are fan-triangulated; negative indices count
    back from the last vertex read; texture/normal indices are ignored."""
    verts, tris, name = [], [], None
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for lineno, line in enumerate(f, 1):
            tok = line.split("#", 1)[0].split()
            if not tok:
                continue
            if tok[0] == "v":
                if len(tok) < 4:
                    raise ValueError(f"{path}:{lineno}: vertex needs 3 coordinates")
                try:
                    verts.append([float(tok[1]), float(tok[2]), float(tok[3])])
                except ValueError:
                    raise ValueError(f"{path}:{lineno}: bad vertex coordinate") from None
            elif tok[0] == "f":
                if len(tok) < 4:
                    raise ValueError(f"{path}:{lineno}: face needs 3 corners")
                idx = []
                for t in tok[1:]:
                    try:
                        raw = int(t.split("/", 1)[0])
                    except ValueError:
                        raise ValueError(f"{path}:{lineno}: bad face index {t!r}") from None
                    k = raw - 1 if raw > 0 else len(verts) + raw
                    if raw == 0 or k < 0 or k >= len(verts):
                        raise ValueError(f"{path}:{lineno}: face index {t} out of range")
                    idx.append(k)
                for c in range(1, len(idx) - 1):
                    tris.append([idx[0], idx[c], idx[c + 1]])
            elif tok[0] == "o" and name is None and len(tok) > 1:
                name = " ".join(tok[1:])
    if name is None:
        name = Path(path).stem
    return (np.asarray(verts, dtype=np.float64).reshape(-1, 3),
            np.asarray(tris, dtype=np.int32).reshape(-1, 3), name)


def _f32(x):
    return float(np.float32(x))


def blender_to_solver(p):
    """(x, y, z) in Blender's z-up frame -> (x, z, -y) (seed_export.py:142-147)."""
    p = np.asarray(p, dtype=np.float64)
    return np.stack([p[..., 0], p[..., 2], -p[..., 1]], axis=-1)


class Group:
    """One snow group: a closed mesh and the add-on's per-object settings."""

    def __init__(self, verts, tris, name, velocity=DEFAULT_VELOCITY, mass=DEFAULT_MASS, spacing=DEFAULT_SPACING):
        # the add-on's properties clamp to their minimum (seed_export.py:30-38)
        if not (np.isfinite(mass) and np.isfinite(spacing)):
            raise ValueError("mass and spacing must be finite")
        mass, spacing = max(mass, 0.000001), max(spacing, 0.0001)
        self.verts = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
        self.tris = np.asarray(tris, dtype=np.int32).reshape(-1, 3)
        self.name = name
        self.velocity = tuple(_f32(c) for c in velocity)
        self.mass = _f32(mass)
        self.spacing = _f32(spacing)

    @classmethod
    def from_obj(cls, path, **settings):
        v, t, name = read_obj(path)
        return cls(v, t, name, **settings)


def seed_groups(groups, output_directory, frame="blender", device=0, log=print):
    """Seed every group and write groups.csv and particles.csv into
    output_directory (export_groups, export_particles and the operator's
    execute, seed_export.py:165-317).  Returns the total particle count."""
    if frame not in ("blender", "solver"):
        raise ValueError("frame must be 'blender' or 'solver'")
    if not groups:
        raise ValueError("No enabled snow mesh objects selected")
    groups = sorted(groups, key=lambda g: g.name)
    conv = blender_to_solver if frame == "blender" else (lambda p: np.asarray(p, dtype=np.float64))
    out = Path(output_directory)
    out.mkdir(parents=True, exist_ok=True)
    with (out / "groups.csv").open("w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(["group_id", "vx", "vy", "vz", "mass"])
        for gid, g in enumerate(groups):
            v = conv(np.asarray(g.velocity))
            w.writerow([gid, float(v[0]), float(v[1]), float(v[2]), g.mass])
    total = 0
    with (out / "particles.csv").open("w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(["group_id", "x", "y", "z"])
        for gid, g in enumerate(groups):
            pts = conv(seed_mesh(g.verts, g.tris, g.spacing, device=device))
            for p in pts:
                w.writerow([gid, float(p[0]), float(p[1]), float(p[2])])
            total += len(pts)
            log(f"Group {gid} '{g.name}': {len(pts)} particles")
    log(f"Exported {len(groups)} groups with {total} particles")
    return total


def _parse_group(spec):
    path, settings = spec[0], {}
    for kv in spec[1:]:
        key, _, val = kv.partition("=")
        if key == "velocity":
            comps = val.split(",")
            if len(comps) != 3:
                raise ValueError(f"velocity needs 3 components: {kv}")
            settings["velocity"] = tuple(float(c) for c in comps)
        elif key in ("mass", "spacing"):
            settings[key] = float(val)
        else:
            raise ValueError(f"unknown group setting {kv!r} (velocity=, mass=, spacing=)")
    return Group.from_obj(path, **settings)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m snowmpm.seed", description=__doc__.split("\n\n")[0])
    ap.add_argument("outdir", help="directory for groups.csv and particles.csv")
    ap.add_argument("--group", action="append", nargs="+", metavar="ARG", default=[],
                    help="MESH.obj [velocity=VX,VY,VZ] [mass=M] [spacing=S]")
    ap.add_argument("--frame", choices=("blender", "solver"), default="blender",
                    help="frame of the mesh coordinates (blender: z up, converted to the solver's y up)")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    try:
        groups = [_parse_group(g) for g in a.group]
        seed_groups(groups, a.outdir, frame=a.frame, device=a.device)
    except (ValueError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
