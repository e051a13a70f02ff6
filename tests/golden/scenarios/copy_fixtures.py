"""Copies the reference's scenario files used as fixtures into this folder and
records their SHA-256 in MANIFEST.txt.  Run in a container holding the
reference tree at /root/reference:

    python tests/golden/scenarios/copy_fixtures.py

The .data files are LZ4 frames written by the reference's ParticleSerializer;
the .json files are its scenario files (version 2).  Both are data: inputs and
the reference's own saved outputs, used to check the readers and as simulation
inputs."""
import hashlib
import os
import shutil
import sys

SRC = "/root/reference/scenarios"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = [
    "3d/3d-small-breaking-dam.json",
    "3d/3d-small-breaking-dam.data",
    "3d/boundary-test-multi-layer-3d.json",
    "3d/boundary-test-multi-layer-3d.data",
    "3d/single-layer/filled-container.json",
    "3d/single-layer/filled-container.data",
    "3d/single-layer/breaking-dam-1-layer-boundary.json",
    "3d/single-layer/breaking-dam-1-layer-boundary.data",
    "3d/3d-100x100x100-breaking-dam.json",
    "3d/single-layer/bunny-teapot.json",
    "2d/box.json",
    "2d/box.data",
    "2d/boundary-test.json",
    "2d/boundary-test.data",
]


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} is not present")
    lines = []
    for rel in FILES:
        dst = os.path.join(HERE, os.path.basename(rel))  # basenames are unique; the JSON files name them
        shutil.copyfile(os.path.join(SRC, rel), dst)
        digest = hashlib.sha256(open(dst, "rb").read()).hexdigest()
        lines.append(f"{digest}  {os.path.basename(dst)}  (scenarios/{rel})")
    with open(os.path.join(HERE, "MANIFEST.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
