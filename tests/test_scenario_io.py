"""Scenario files on the host (no GPU): xxHash32, the LZ4 frame codec, the
particle .data reader/writer and the JSON version-2 scenario reader/writer.

Pins:
  - xxh32 against the xxhash package and the published answers for "" and "abc";
  - the LZ4 decoders (library and oracle) against hand-built frames whose
    content follows from the format definition (overlapping matches, length
    extension bytes, linked and independent blocks, checksums, skippable
    frames, and the malformed cases the format forbids);
  - the particle reader against the reference's own scenario files, whose
    particle counts and fluid/boundary split were read independently with the
    lz4 command-line tool (mass = density = 1 in all of them);
  - the scenario reader against Python's json module on the reference's
    scenario files (ScenarioSerializer / SolverSerializer / TimestepGeneratorSerializer).
"""
import json
import os
import struct

import numpy as np
import pytest

from fluidsolver_amd import abi
from fluidsolver_amd import scenario as S
from fluidsolver_amd.collection import ParticleCollection
from fluidsolver_amd.solver import FsaError
from oracle import scenario_io as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "scenarios")

# (file, particles, fluid, boundary, dimension, component ids)
FIXTURES = [
    ("3d-small-breaking-dam.data", 5764, 1000, 4764, 3, [2, 3, 4, 6]),
    ("filled-container.data", 3656, 2535, 1121, 3, [2, 3, 4, 6]),
    ("boundary-test-multi-layer-3d.data", 364, 1, 363, 3, [2, 3, 4, 6]),
    ("box.data", 753, 441, 312, 2, [1, 3, 4, 5]),
    ("boundary-test.data", 64, 1, 63, 2, [1, 3, 4, 5]),
]


# ---- xxHash32 -------------------------------------------------------------------------
def test_xxh32_known_answers():
    assert O.xxh32(b"") == 0x02CC5D05
    assert O.xxh32(b"abc") == 0x32D153FF
    assert S.xxh32(b"") == 0x02CC5D05
    assert S.xxh32(b"abc") == 0x32D153FF


def test_xxh32_against_xxhash_package():
    xxhash = pytest.importorskip("xxhash")
    rng = np.random.default_rng(5)
    for n in list(range(0, 70)) + [255, 256, 1000, 4097]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        for seed in (0, 1, 0x9E3779B1):
            want = xxhash.xxh32(data, seed=seed).intdigest()
            assert S.xxh32(data, seed) == want, (n, seed)
            if n <= 300:
                assert O.xxh32(data, seed) == want, (n, seed)


# ---- LZ4 frames -------------------------------------------------------------------------
def frame(blocks, flg=0x40, bd=0x40, content=None, content_size=None):
    """An LZ4 frame from (raw, bytes) blocks; flg bits: 0x20 independent,
    0x10 block checksums, 0x08 content size, 0x04 content checksum."""
    desc = bytes([flg, bd]) + (struct.pack("<Q", content_size) if flg & 0x08 else b"")
    out = struct.pack("<I", 0x184D2204) + desc + bytes([(O.xxh32(desc) >> 8) & 0xFF])
    for raw, data in blocks:
        out += struct.pack("<I", len(data) | (0x80000000 if raw else 0)) + data
        if flg & 0x10:
            out += struct.pack("<I", O.xxh32(data))
    out += struct.pack("<I", 0)
    if flg & 0x04:
        out += struct.pack("<I", O.xxh32(content))
    return out


def both_decode(data):
    a = S.lz4_frame_decode(data)
    b = O.lz4_frame_decode(data)
    assert a == b
    return a


def both_reject(data):
    with pytest.raises(FsaError) as e:
        S.lz4_frame_decode(data)
    assert e.value.code == abi.FSA_E_IO
    with pytest.raises((O.FormatError, IndexError, struct.error)):
        O.lz4_frame_decode(data)


def test_lz4_uncompressed_block():
    assert both_decode(frame([(True, b"Hello, particles")])) == b"Hello, particles"


def test_lz4_overlapping_match_and_last_literals():
    # token: 2 literals, match length 4 + 6; offset 2 repeats "ab"; then 5 literals
    block = bytes([0x26]) + b"ab" + bytes([2, 0]) + bytes([0x50]) + b"cdefg"
    assert both_decode(frame([(False, block)])) == b"ab" * 6 + b"cdefg"


def test_lz4_length_extension_bytes():
    lits = bytes(range(256)) + bytes(range(44))  # 300 literals: 15 + 255 + 30
    block = bytes([0xF0, 255, 30]) + lits
    assert both_decode(frame([(False, block)])) == lits
    # match length 4 + 15 + 255 + 1 = 275 at offset 1 (a run), then 5 literals
    block = bytes([0x1F]) + b"z" + bytes([1, 0, 255, 1]) + bytes([0x50]) + b"12345"
    assert both_decode(frame([(False, block)])) == b"z" * 276 + b"12345"


def test_lz4_linked_blocks_reach_back():
    first = b"0123456789abcdef"
    # second block: 0 literals, match at offset 16 (into the first block) of length 8, then 5 literals
    second = bytes([0x04]) + bytes([16, 0]) + bytes([0x50]) + b"XYZWV"
    linked = frame([(True, first), (False, second)], flg=0x40)
    assert both_decode(linked) == first + b"01234567" + b"XYZWV"
    both_reject(frame([(True, first), (False, second)], flg=0x60))  # independent blocks may not


def test_lz4_checksums_sizes_and_skippable_frames():
    content = b"checksummed content" * 3
    good = frame([(True, content)], flg=0x40 | 0x10 | 0x08 | 0x04, content=content, content_size=len(content))
    assert both_decode(good) == content
    skip = struct.pack("<II", 0x184D2A5F, 3) + b"xyz"
    assert both_decode(skip + good) == content
    bad = bytearray(good)
    bad[-1] ^= 1  # content checksum
    both_reject(bytes(bad))
    bad = bytearray(good)
    bad[6] ^= 1  # inside the content size: the header checksum no longer matches
    both_reject(bytes(bad))
    both_reject(frame([(True, content)], flg=0x48, content_size=len(content) + 1))


def test_lz4_malformed_blocks():
    both_reject(frame([(False, bytes([0x14]) + b"a" + bytes([2, 0]))]))  # offset before the output start
    both_reject(frame([(False, bytes([0x10]) + b"a" + bytes([0, 0]) + bytes([0x00]))]))  # offset 0
    both_reject(frame([(False, bytes([0x50]) + b"abc")]))  # literals past the block
    both_reject(b"\x00\x01\x02\x03garbage")
    both_reject(frame([(True, b"x")])[:-2])  # truncated


@pytest.mark.parametrize("kind", ["random", "runs", "zeros", "records"])
@pytest.mark.parametrize("n", [0, 1, 12, 13, 100, 65535, 65536, 65537, 200003])
def test_lz4_encoder_round_trip(kind, n):
    rng = np.random.default_rng(n)
    if kind == "random":
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
    elif kind == "runs":
        data = np.repeat(rng.integers(0, 4, n // 7 + 1, dtype=np.uint8), 7)[:n].tobytes()
    elif kind == "zeros":
        data = bytes(n)
    else:
        data = np.tile(np.arange(65, dtype=np.uint8), n // 65 + 1)[:n].tobytes()
    enc = S.lz4_frame_encode(data)
    assert S.lz4_frame_decode(enc) == data
    if n <= 70000:
        assert O.lz4_frame_decode(enc) == data
    if n >= 1000:  # compressible inputs shrink (runs of 7 equal bytes carry about 2 bits)
        bound = {"zeros": n // 50, "records": n // 20, "runs": n // 2, "random": n + n // 100 + 64}[kind]
        assert len(enc) < bound


# ---- particle files ----------------------------------------------------------------------------
@pytest.mark.parametrize("name,n,nf,nb,dim,ids", FIXTURES)
def test_reference_particle_files(name, n, nf, nb, dim, ids):
    path = os.path.join(GOLD, name)
    version, fids, rec = O.parse_particle_file(O.lz4_frame_decode(open(path, "rb").read()))
    assert version == 1 and fids == ids and len(rec) == n
    h = S.read_particle_header(path)
    assert (h.version, h.dimension, h.count, h.component_ids()) == (1, dim, n, ids)
    coll = S.read_particles(path)
    assert coll.dimension == dim and coll.size() == n
    counts = coll.counts()
    assert (counts["fluid"], counts["boundary"]) == (nf, nb)
    assert np.all(coll.mass == 1.0) and np.all(coll.density == 1.0)
    # field by field against the oracle's parse
    np.testing.assert_array_equal(coll.position, rec["position"])
    np.testing.assert_array_equal(coll.velocity, rec["velocity"])
    np.testing.assert_array_equal(coll.acceleration, rec["acceleration"])
    np.testing.assert_array_equal(coll.type, rec["type"])
    np.testing.assert_array_equal(coll.tag, rec["tag"])
    np.testing.assert_array_equal(coll.pressure, rec["pressure"])
    np.testing.assert_array_equal(coll.external, rec["external"])


@pytest.mark.parametrize("name", [f[0] for f in FIXTURES])
def test_particle_file_round_trip(tmp_path, name):
    src = os.path.join(GOLD, name)
    coll = S.read_particles(src)
    out = tmp_path / "copy.data"
    S.write_particles(out, coll)
    again = S.read_particles(out)
    assert S.records_equal(coll, again)
    # the written records are byte-identical to the reference's
    a = O.parse_particle_file(O.lz4_frame_decode(open(src, "rb").read()))[2]
    b = O.parse_particle_file(O.lz4_frame_decode(open(out, "rb").read()))[2]
    assert a.tobytes() == b.tobytes()


def test_particle_file_components_and_views(tmp_path):
    rng = np.random.default_rng(3)
    coll = ParticleCollection(3, 50)
    coll.movement[:] = rng.standard_normal(coll.movement.shape).astype(np.float32)
    coll.particle_data[:] = rng.random(coll.particle_data.shape).astype(np.float32)
    coll.info["tag"] = rng.integers(0, 2**32, 50, dtype=np.uint64).astype(np.uint32)
    coll.info["type"] = rng.integers(0, 3, 50).astype(np.uint8)
    coll.external[:] = rng.standard_normal(coll.external.shape).astype(np.float32)
    p = tmp_path / "noext.data"
    S.write_particles(p, coll, external=False)
    assert S.read_particle_header(p).component_ids() == [2, 3, 4]
    back = S.read_particles(p)
    assert np.all(back.external == 0.0)
    np.testing.assert_array_equal(back.movement, coll.movement)
    np.testing.assert_array_equal(back.info, coll.info)
    # a view whose count differs from the file is refused
    bad = abi.make_view(ParticleCollection(3, 49))
    from fluidsolver_amd._native import lib
    import ctypes as C
    assert lib().fsa_data_read(os.fsencode(str(p)), C.byref(bad)) == abi.FSA_E_INVALID


def test_particle_file_errors(tmp_path):
    content = struct.pack("<III", 2, 1, 2) + struct.pack("<Q", 0)
    p = tmp_path / "v2.data"
    p.write_bytes(S.lz4_frame_encode(content))
    with pytest.raises(FsaError, match="Unsupported particle data format version"):
        S.read_particles(p)
    content = struct.pack("<III", 1, 1, 9) + struct.pack("<Q", 0)
    p.write_bytes(S.lz4_frame_encode(content))
    with pytest.raises(FsaError, match="invalid component id"):
        S.read_particles(p)
    content = struct.pack("<III", 1, 1, 2) + struct.pack("<Q", 3) + bytes(36 * 2)
    p.write_bytes(S.lz4_frame_encode(content))
    with pytest.raises(FsaError, match="fewer records"):
        S.read_particles(p)
    with pytest.raises(FsaError, match="cannot open"):
        S.read_particles(tmp_path / "missing.data")


# ---- scenario files ------------------------------------------------------------------------------
SOLVERS = {
    ("sesph", "quadratic-dynamic-allocated"): (2, abi.SOLVER_SESPH, abi.SEARCH_QUADRATIC_2D),
    ("sesph", "hashed"): (2, abi.SOLVER_SESPH, abi.SEARCH_HASHED_2D),
    ("iisph", "quadratic-dynamic-allocated"): (2, abi.SOLVER_IISPH, abi.SEARCH_QUADRATIC_2D),
    ("iisph", "hashed"): (2, abi.SOLVER_IISPH, abi.SEARCH_HASHED_2D),
    ("sesph-3d", "hashed-3d"): (3, abi.SOLVER_SESPH, abi.SEARCH_HASHED_3D),
    ("sesph-3d", "compressed-3d"): (3, abi.SOLVER_SESPH, abi.SEARCH_COMPRESSED_3D),
    ("sesph-3d", "quadratic-dynamic-allocated-3d"): (3, abi.SOLVER_SESPH, abi.SEARCH_QUADRATIC_3D),
    ("iisph-3d", "hashed-3d"): (3, abi.SOLVER_IISPH, abi.SEARCH_HASHED_3D),
    ("iisph-3d", "compressed-3d"): (3, abi.SOLVER_IISPH, abi.SEARCH_COMPRESSED_3D),
    ("iisph-3d", "quadratic-dynamic-allocated-3d"): (3, abi.SOLVER_IISPH, abi.SEARCH_QUADRATIC_3D),
}
f32 = lambda x: float(np.float32(x))  # noqa: E731


def check_against_json(sc, j):
    d = sc.desc
    so, scn, tg = j["solver"], j["scenario"], j["timestep-generator"]
    dim, solver, search = SOLVERS[(so["type"], so["neighborhood-search"]["type"])]
    assert (d.dimension, d.solver, d.search) == (dim, solver, search)
    assert sc.particles == scn["particles"]
    assert (d.gravity, d.particle_size, d.rest_density) == (
        f32(scn["gravity"]), f32(scn["particle-size"]), f32(scn["rest-density"]))
    if solver == abi.SOLVER_SESPH:
        assert (d.sesph_stiffness_k, d.sesph_viscosity) == (f32(so["stiffness"]), f32(so["viscosity"]))
        if dim == 3:
            assert bool(d.sesph_single_layer_boundary) == so["single-layer-boundary-enabled"]
            if so["single-layer-boundary-enabled"]:
                assert d.sesph_single_layer_gamma_1 == f32(so["single-layer-settings"]["gamma-1"])
                assert d.sesph_single_layer_gamma_2 == f32(so["single-layer-settings"]["gamma-2"])
    else:
        assert (d.iisph_gamma, d.iisph_omega, d.iisph_max_density_error_allowed, d.iisph_viscosity) == (
            f32(so["gamma"]), f32(so["omega"]), f32(so["max-density-error"]), f32(so["viscosity"]))
        assert (d.iisph_max_iterations, d.iisph_min_iterations) == (so["max-iterations"], so["min-iterations"])
        if dim == 3:
            assert bool(d.iisph_single_layer_boundary) == so["single-layer-boundary-enabled"]
            if so["single-layer-boundary-enabled"]:
                assert d.iisph_single_layer_gamma_1 == f32(so["single-layer-settings"]["gamma-1"])
    if tg["type"] == "constant":
        assert sc.timestep_generator == abi.TIMESTEP_CONSTANT and d.timestep == f32(tg["timestep"])
    else:
        assert sc.dynamic_cfl == {"max_timestep": f32(tg["max-timestep"]), "min_timestep": f32(tg["min-timestep"]),
                                  "cfl": f32(tg["cfl"])}
    assert sc.entities == len(scn.get("entities", [])) and sc.sensors == len(scn.get("sensors", []))


JSONS = sorted(f for f in os.listdir(GOLD) if f.endswith(".json"))


@pytest.mark.parametrize("name", JSONS)
def test_reference_scenario_files(name):
    path = os.path.join(GOLD, name)
    sc = S.load_scenario(path, load_particles=False)
    check_against_json(sc, json.load(open(path)))
    if os.path.exists(os.path.join(GOLD, sc.particles)):
        assert sc.particles_path == os.path.join(GOLD, sc.particles)
        full = S.load_scenario(path)
        assert full.collection.dimension == sc.desc.dimension


@pytest.mark.parametrize("name", JSONS)
def test_scenario_round_trip(tmp_path, name):
    src = S.load_scenario(os.path.join(GOLD, name), load_particles=False)
    out = tmp_path / "again.json"
    S.save_scenario(out, src)
    j = json.load(open(out))
    assert j["version"] == 2 and j["visualizer"] == {"type": "no-visualizer"}
    again = S.load_scenario(out, load_particles=False)
    check_against_json(again, j)
    a, b = bytes(src.raw.desc), bytes(again.raw.desc)
    assert a == b


def test_scenario_save_with_particles_and_dynamic_cfl(tmp_path):
    sc = S.load_scenario(os.path.join(GOLD, "3d-small-breaking-dam.json"))
    d = sc.raw.desc
    d.timestep_generator = abi.TIMESTEP_DYNAMIC_CFL
    d.max_timestep, d.min_timestep, d.cfl_number = 0.01, 0.0005, 0.7
    out = tmp_path / "dump" / "snap.json"
    out.parent.mkdir()
    S.save_scenario(out, sc, particles="snap.data")
    back = S.load_scenario(out)
    assert back.particles_path == str(tmp_path / "dump" / "snap.data")
    assert S.records_equal(back.collection, sc.collection)
    assert back.dynamic_cfl == {"max_timestep": f32(0.01), "min_timestep": f32(0.0005), "cfl": f32(0.7)}


def test_scenario_errors(tmp_path):
    base = json.load(open(os.path.join(GOLD, "3d-small-breaking-dam.json")))

    def load(obj, text=None):
        p = tmp_path / "s.json"
        p.write_text(text if text is not None else json.dumps(obj))
        return S.load_scenario(p, load_particles=False)

    j = json.loads(json.dumps(base))
    j["version"] = 1
    with pytest.raises(FsaError, match="version"):
        load(j)
    j = json.loads(json.dumps(base))
    j["solver"]["neighborhood-search"]["type"] = "hashed"  # a 2D search under a 3D solver
    with pytest.raises(FsaError, match="unknown solver, neighborhood search, kernel combination"):
        load(j)
    j = json.loads(json.dumps(base))
    del j["solver"]["stiffness"]
    with pytest.raises(FsaError, match="solver.stiffness"):
        load(j)
    j = json.loads(json.dumps(base))
    j["timestep-generator"]["type"] = "adaptive"
    with pytest.raises(FsaError, match="timestep generator"):
        load(j)
    with pytest.raises(FsaError, match="JSON"):
        load(None, text=json.dumps(base)[:-3])
    with pytest.raises(FsaError, match="JSON"):
        load(None, text="[" * 500 + "]" * 500)
    # escapes and unicode in strings survive
    j = json.loads(json.dumps(base))
    j["scenario"]["particles"] = "dépôt/\"x\"\\€\U0001F30A.data"
    assert load(j).particles == j["scenario"]["particles"]


# ---- entities in scenario files (EntitySerializer.cpp:66-173) ------------------------------------
def test_scenario_entities_from_reference_file():
    path = os.path.join(GOLD, "bunny-teapot.json")
    j = json.load(open(path))
    sc = S.load_scenario(path, load_particles=False)
    ents = j["scenario"]["entities"]
    assert sc.entities == len(ents)
    runnable = [e for e in ents if e["type"] in ("particle-remover-3d", "boundary-preprocessor",
                                                 "velocity-alteration-by-tag")]
    assert len(sc.entity_list) == len(runnable) >= 1
    for got, e in zip(sc.entity_list, runnable):
        assert got.type == abi.ENTITY_PARTICLE_REMOVER_3D
        assert list(got.center) == [f32(x) for x in e["volume"]["center"]]
        assert list(got.distance_from_center) == [f32(x) for x in e["volume"]["distance-from-center"]]
        assert got.remove_if_outside == int(e["remove-if-outside"])


def test_scenario_entities_round_trip(tmp_path):
    sc = S.load_scenario(os.path.join(GOLD, "3d-small-breaking-dam.json"), load_particles=False)
    sc.set_entity_list([abi.boundary_preprocessor(), abi.velocity_alteration_by_tag(7, [0.5, -1.25, 2.0]),
                        abi.particle_remover_3d([1.0, 2.0, 3.0], [4.0, 5.5, 6.25], True)])
    out = tmp_path / "ents.json"
    S.save_scenario(out, sc)
    j = json.load(open(out))
    assert [e["type"] for e in j["scenario"]["entities"]] == [
        "boundary-preprocessor", "velocity-alteration-by-tag", "particle-remover-3d"]
    back = S.load_scenario(out, load_particles=False)
    assert [bytes(e) for e in back.entity_list] == [bytes(e) for e in sc.entity_list]
    # unknown entity types are errors, 2D ones are counted only
    j["scenario"]["entities"].append({"type": "particle-spawner"})
    out.write_text(json.dumps(j))
    back = S.load_scenario(out, load_particles=False)
    assert back.entities == 4 and len(back.entity_list) == 3
    j["scenario"]["entities"].append({"type": "teleporter"})
    out.write_text(json.dumps(j))
    with pytest.raises(FsaError, match="invalid entity type"):
        S.load_scenario(out, load_particles=False)
