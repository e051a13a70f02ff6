"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, and the ctypes mirrors match the C struct layouts.  No device calls."""
import ctypes as C
import os
import subprocess

import pytest

from fluidsolver_amd import _native, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    syms = _native.header_symbols()
    assert len(syms) >= 25
    L = _native.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []
    # nothing beyond the header is exported with the fsa_ prefix
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if ln.split()[-1].startswith("fsa_")})
    assert exported == syms


def test_version_and_defaults_without_device():
    L = _native.lib()
    assert L.fsa_version() >= 1
    d = abi.SceneDesc()
    L.fsa_scene_desc_defaults(C.byref(d), 3, abi.SOLVER_IISPH)
    ref = abi.default_desc(3, abi.SOLVER_IISPH)
    for name, _ in abi.SceneDesc._fields_:
        if name in ("neighbor_lists",):
            continue
        assert getattr(d, name) == getattr(ref, name), name


def test_memory_estimate_and_invalid_desc():
    L = _native.lib()
    d = abi.default_desc(3, abi.SOLVER_IISPH)
    v = C.c_uint64(0)
    assert L.fsa_scene_memory_estimate(C.byref(d), 1000000, 696816, C.byref(v)) == abi.FSA_OK
    assert 300e6 < v.value < 2e9
    bad = abi.default_desc(3, abi.SOLVER_IISPH, iisph_max_iterations=1, iisph_min_iterations=2)
    assert L.fsa_scene_memory_estimate(C.byref(bad), 10, 10, C.byref(v)) == abi.FSA_E_INVALID
    assert b"Max iterations" in L.fsa_last_error()
    bad = abi.default_desc(3, abi.SOLVER_IISPH, particle_size=0.0)
    assert L.fsa_scene_memory_estimate(C.byref(bad), 10, 10, C.byref(v)) == abi.FSA_E_INVALID


def test_null_arguments_return_status():
    L = _native.lib()
    assert L.fsa_scene_find_neighbors(None) == abi.FSA_E_INVALID
    assert L.fsa_scene_simulate(None, 0.001, None) == abi.FSA_E_INVALID
    assert L.fsa_context_create(0, 0, None) == abi.FSA_E_INVALID
    assert L.fsa_context_set_walk_variant(None, abi.WALK_U4) == abi.FSA_E_INVALID
    assert L.fsa_server_create(None, None, None) == abi.FSA_E_INVALID
    assert L.fsa_server_tick(None, None) == abi.FSA_E_INVALID
    assert L.fsa_server_submit(None, None, None, 1, None) == abi.FSA_E_INVALID
    assert L.fsa_server_get_stats(None, None) == abi.FSA_E_INVALID
    nnz = C.c_uint64(0)
    assert L.fsa_scene_query_neighbors(None, None, 0, None, None, 0, C.byref(nnz)) == abi.FSA_E_INVALID
    assert L.fsa_scene_sensor_plane(None, None, None) == abi.FSA_E_INVALID
    assert L.fsa_scene_evaluate_volume(None, None, 0, None) == abi.FSA_E_INVALID


def test_sensor_plane_defaults_match_reference():
    """SensorPlaneSettings defaults (SensorPlane.hpp:17-37), from the library and the Python builder."""
    p = abi.SensorPlane()
    _native.lib().fsa_sensor_plane_defaults(C.byref(p))
    assert bytes(p) == bytes(abi.sensor_plane())
    assert (p.width, p.height, p.number_of_samples_x, p.number_of_samples_y, p.sub_sample_grid_size) == (
        30.0, 10.0, 300, 100, 2)
    assert list(p.span_x) == [1.0, 0.0, 0.0] and list(p.span_y) == [0.0, 1.0, 0.0]


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "fluidsolver_amd.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(fsa_scene_desc), sizeof(fsa_particle_view), sizeof(fsa_step_stats));
  printf("%zu %zu %zu\n", offsetof(fsa_scene_desc, timestep), offsetof(fsa_particle_view, external_stride),
         offsetof(fsa_step_stats, average_density_error));
  printf("%zu %zu %zu\n", sizeof(fsa_server_config), sizeof(fsa_job_info), sizeof(fsa_server_stats));
  printf("%zu %zu %zu\n", offsetof(fsa_job_info, last), offsetof(fsa_server_stats, busy_s),
         offsetof(fsa_server_stats, max_step_ms));
  printf("%zu %zu %zu\n", sizeof(fsa_sensor_plane), offsetof(fsa_sensor_plane, sub_sample_grid_size),
         sizeof(fsa_volume_sample));
  return 0;
}
"""


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes[0] == C.sizeof(abi.SceneDesc)
    assert sizes[1] == C.sizeof(abi.ParticleView)
    assert sizes[2] == C.sizeof(abi.StepStats)
    assert sizes[3] == abi.SceneDesc.timestep.offset
    assert sizes[4] == abi.ParticleView.external_stride.offset
    assert sizes[5] == abi.StepStats.average_density_error.offset
    assert sizes[6:9] == [C.sizeof(abi.ServerConfig), C.sizeof(abi.JobInfo), C.sizeof(abi.ServerStats)]
    assert sizes[9] == abi.JobInfo.last.offset
    assert sizes[10] == abi.ServerStats.busy_s.offset
    assert sizes[11] == abi.ServerStats.max_step_ms.offset
    assert sizes[12] == C.sizeof(abi.SensorPlane)
    assert sizes[13] == abi.SensorPlane.sub_sample_grid_size.offset
    assert sizes[14] == C.sizeof(abi.VolumeSample)


def test_info_record_layout():
    """ParticleInfo {uint32 tag; uint8 type} is 8 bytes with type at offset 4 (Particle.hpp:40-44)."""
    assert abi.INFO_DTYPE.itemsize == 8
    assert abi.INFO_DTYPE.fields["type"][1] == 4


@pytest.mark.parametrize("dim", [2, 3])
def test_collection_views_are_zero_copy(dim):
    from fluidsolver_amd.collection import ParticleCollection

    c = ParticleCollection(dim, 7)
    v = abi.make_view(c)
    assert v.count == 7
    assert v.movement == c.movement.ctypes.data
    assert v.movement_stride == 4 * 3 * dim
    assert v.particle_data_stride == 12
    assert v.info_stride == 8
    assert v.external_stride == 4 * dim
