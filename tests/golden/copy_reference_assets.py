#!/usr/bin/env python3
"""Copies the reference's textured vehicle (CCR/Resources/vehicle.obj and its
diffuse, normal, specular and gloss PNG maps, about 6 MB) into
tests/golden/assets/, which git ignores: the files are data the reference
holds, too large to keep in history.  The GPU test that renders them skips
when they are absent; run this once where the reference is checked out.

    python tests/golden/copy_reference_assets.py [/path/to/reference]
"""
import os
import shutil
import sys

FILES = ["vehicle.obj", "vehicle_diffuse.png", "vehicle_normal.png", "vehicle_specular.png", "vehicle_gloss.png"]
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "CustomCUDARasterizer", "Resources")
    dst = os.path.join(HERE, "assets")
    os.makedirs(dst, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        print("copied", f)


if __name__ == "__main__":
    main()
