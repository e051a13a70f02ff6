"""Writes tests/golden/*.npz from the reference's own scenario fixtures.

Run where the reference checkout is present (it reads /root/reference; the
machines that run the GPU tests only read the committed fixtures):

    python tests/golden/make_fixtures.py

Each `.data` file is an LZ4 frame (decoded here with the `unlz4` tool) holding
the ParticleSerializer format (ParticleSerializer.cpp:41-141): u32 version = 1,
u32 component count, u32 component ids, u64 particle count, then per particle
the present components in this order: MovementData (2D) / MovementData3D
(position, velocity, acceleration), ParticleInfo (u8 type, u32 tag),
ParticleData (density, mass, pressure), ExternalForces / ExternalForces3D; all
little-endian and unpadded.  The fixture stores them in this repository's
component layouts plus the solver settings of the scenario's JSON.
"""
import json
import os
import struct
import subprocess
import sys

import numpy as np

REF = "/root/reference/scenarios"
HERE = os.path.dirname(os.path.abspath(__file__))
UNLZ4 = "/opt/conda/bin/unlz4"

FIXTURES = {
    "small_breaking_dam_3d": "3d/3d-small-breaking-dam",
    "filled_container_single_layer_3d": "3d/single-layer/filled-container",
    "boundary_test_multi_layer_3d": "3d/boundary-test-multi-layer-3d",
    "box_2d": "2d/box",
    "boundary_test_2d": "2d/boundary-test",
}

COMP_MOVEMENT, COMP_MOVEMENT3D, COMP_INFO, COMP_PDATA, COMP_EXT, COMP_EXT3D = 1, 2, 3, 4, 5, 6


def decode(path):
    raw = subprocess.run([UNLZ4, "-c", "-q", path], capture_output=True, check=True).stdout
    off = 0
    version, ncomp = struct.unpack_from("<II", raw, off)
    off += 8
    if version != 1:
        raise ValueError(f"unsupported particle data version {version}")
    ids = set(struct.unpack_from(f"<{ncomp}I", raw, off))
    off += 4 * ncomp
    (n,) = struct.unpack_from("<Q", raw, off)
    off += 8
    dim = 3 if COMP_MOVEMENT3D in ids else 2
    fields = []
    if COMP_MOVEMENT in ids:
        fields.append(("movement", "<f4", (6,)))
    if COMP_MOVEMENT3D in ids:
        fields.append(("movement", "<f4", (9,)))
    if COMP_INFO in ids:
        fields += [("type", "u1"), ("tag", "<u4")]
    if COMP_PDATA in ids:
        fields += [("density", "<f4"), ("mass", "<f4"), ("pressure", "<f4")]
    if COMP_EXT in ids:
        fields.append(("external", "<f4", (2,)))
    if COMP_EXT3D in ids:
        fields.append(("external", "<f4", (3,)))
    rec = np.frombuffer(raw, dtype=np.dtype(fields), count=n, offset=off)
    if off + rec.nbytes != len(raw):
        raise ValueError("trailing bytes in particle data")
    out = {
        "dimension": np.int32(dim),
        "movement": np.ascontiguousarray(rec["movement"], dtype=np.float32),
        "particle_data": np.stack([rec["mass"], rec["pressure"], rec["density"]], 1).astype(np.float32),
        "tag": rec["tag"].astype(np.uint32),
        "type": rec["type"].astype(np.uint8),
    }
    out["external"] = (
        np.ascontiguousarray(rec["external"], dtype=np.float32) if "external" in rec.dtype.names
        else np.zeros((n, dim), np.float32)
    )
    return out


def settings(jpath):
    j = json.load(open(jpath))
    sc, so, tg = j["scenario"], j["solver"], j["timestep-generator"]
    s = {
        "solver": so["type"],
        "search": so["neighborhood-search"]["type"],
        "gravity": sc["gravity"],
        "particle_size": sc["particle-size"],
        "rest_density": sc["rest-density"],
        "timestep": tg["timestep"],
    }
    for k in ("stiffness", "viscosity", "gamma", "omega", "max-density-error", "min-iterations", "max-iterations"):
        if k in so:
            s[k.replace("-", "_")] = so[k]
    s["single_layer"] = bool(so.get("single-layer-boundary-enabled", False))
    sl = so.get("single-layer-settings", {})
    for k in ("gamma-1", "gamma-2"):
        if k in sl:
            s["single_layer_" + k.replace("-", "_")] = sl[k]
    return s


def main():
    if not os.path.isdir(REF):
        print("reference checkout not found; fixtures are committed already", file=sys.stderr)
        return 1
    for name, stem in FIXTURES.items():
        data = decode(os.path.join(REF, stem + ".data"))
        data["settings_json"] = np.frombuffer(json.dumps(settings(os.path.join(REF, stem + ".json"))).encode(), np.uint8)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        t = data["type"]
        print(f"{name}: n={len(t)} fluid={int((t == 0).sum())} boundary={int((t == 1).sum())}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
