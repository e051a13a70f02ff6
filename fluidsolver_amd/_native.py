"""Loader of the in-tree gfx950 library (fluidsolver_amd/lib/libfluidsolver_amd.so).

There is no CPU fallback: every product call goes through this library and a
missing or unloadable library raises immediately.
"""
import ctypes as C
import os
import re

from .abi import (DataHeader, Entity, JobInfo, ParticleView, Scenario, SceneDesc, SceneStatistics, SensorPlane,
                  ServerConfig, ServerStats, StepStats, VolumeSample)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfluidsolver_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fluidsolver_amd.h")

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def _declare(L):
    vp, u32, u64, i32, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_float
    P = C.POINTER
    sig = {
        "fsa_version": (i32, []),
        "fsa_last_error": (C.c_char_p, []),
        "fsa_scene_desc_defaults": (None, [P(SceneDesc), u32, u32]),
        "fsa_context_create": (i32, [i32, u64, P(vp)]),
        "fsa_context_destroy": (i32, [vp]),
        "fsa_context_synchronize": (i32, [vp]),
        "fsa_context_memory_in_use": (i32, [vp, P(u64)]),
        "fsa_scene_memory_estimate": (i32, [P(SceneDesc), u64, u64, P(u64)]),
        "fsa_scene_create": (i32, [vp, P(SceneDesc), P(ParticleView), P(vp)]),
        "fsa_scene_destroy": (i32, [vp]),
        "fsa_scene_upload": (i32, [vp, P(ParticleView)]),
        "fsa_scene_download": (i32, [vp, P(ParticleView)]),
        "fsa_scene_download_iisph": (i32, [vp, vp, u64]),
        "fsa_scene_set_desc": (i32, [vp, P(SceneDesc)]),
        "fsa_scene_find_neighbors": (i32, [vp]),
        "fsa_scene_get_neighbors": (i32, [vp, u64, P(u64), u64, P(u64)]),
        "fsa_scene_export_neighbors": (i32, [vp, vp, vp, u64, P(u64)]),
        "fsa_scene_neighbor_stats": (i32, [vp, P(u64), P(u64), P(u64)]),
        "fsa_scene_simulate": (i32, [vp, f32, P(StepStats)]),
        "fsa_scene_step": (i32, [vp, u32, P(StepStats)]),
        "fsa_scene_next_timestep": (i32, [vp, P(f32)]),
        "fsa_scene_set_entities": (i32, [vp, P(Entity), u32]),
        "fsa_scene_get_statistics": (i32, [vp, f32, P(SceneStatistics)]),
        "fsa_scene_get_statistics_noalloc": (i32, [vp, f32, P(SceneStatistics)]),
        "fsa_scene_query_neighbors": (i32, [vp, vp, u64, vp, vp, u64, P(u64)]),
        "fsa_sensor_plane_defaults": (None, [P(SensorPlane)]),
        "fsa_scene_sensor_plane": (i32, [vp, P(SensorPlane), vp]),
        "fsa_scene_evaluate_volume": (i32, [vp, vp, u64, P(VolumeSample)]),
        "fsa_batch_step": (i32, [vp, P(vp), u32, u32, P(StepStats)]),
        "fsa_batch_set_graph": (i32, [vp, u32]),
        "fsa_context_set_block_size": (i32, [vp, u32]),
        "fsa_context_set_autotune": (i32, [vp, u32]),
        "fsa_context_set_walk_variant": (i32, [vp, u32]),
        "fsa_context_get_tuning": (i32, [vp, C.c_char_p, u64, P(u64)]),
        "fsa_context_set_timing": (i32, [vp, u32]),
        "fsa_context_get_timing": (i32, [vp, P(C.c_double), P(u64), u32]),
        "fsa_server_create": (i32, [vp, P(ServerConfig), P(vp)]),
        "fsa_server_destroy": (i32, [vp]),
        "fsa_server_submit": (i32, [vp, P(SceneDesc), P(ParticleView), u32, P(u64)]),
        "fsa_server_cancel": (i32, [vp, u64]),
        "fsa_server_tick": (i32, [vp, P(u32)]),
        "fsa_server_job": (i32, [vp, u64, P(JobInfo)]),
        "fsa_server_collect": (i32, [vp, u64, P(ParticleView)]),
        "fsa_server_get_stats": (i32, [vp, P(ServerStats)]),
        "fsa_data_read_header": (i32, [C.c_char_p, P(DataHeader)]),
        "fsa_data_read": (i32, [C.c_char_p, P(ParticleView)]),
        "fsa_data_write": (i32, [C.c_char_p, u32, P(ParticleView)]),
        "fsa_scenario_read": (i32, [C.c_char_p, P(Scenario)]),
        "fsa_scenario_write": (i32, [C.c_char_p, P(Scenario)]),
        "fsa_scenario_defaults": (None, [P(Scenario), u32, u32]),
        "fsa_lz4_frame_decode": (i32, [vp, u64, vp, u64, P(u64)]),
        "fsa_lz4_frame_encode": (i32, [vp, u64, vp, u64, P(u64)]),
        "fsa_xxh32": (u32, [vp, u64, u32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    """The loaded device library; raises NativeLibraryMissing when absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
    _lib = _declare(L)
    return _lib


def header_symbols():
    """Every function declared FSA_API in include/fluidsolver_amd.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"FSA_API\s+[\w\s\*]+?\b(fsa_\w+)\s*\(", text)))
