"""Copies the input files the reference ships (input/groups.csv and
input/particles.csv, written by its Blender add-on blender/seed_export.py
from a mesh that is not in the repository) into tests/golden/seed_addon/ as
fixtures for the seeding tool's tests.  Data only: 957 particles of one
group at spacing 0.03, float32 values in Python's repr."""
import os
import shutil

SRC = "/root/reference/input"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seed_addon")

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for name in ("groups.csv", "particles.csv"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
        print("wrote", os.path.join(DST, name))
