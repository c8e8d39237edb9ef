"""ctypes bindings of the two native libraries of the package.

libbwgraph.so       host graph substrate (include/belewitte_graph.h)
libbelewitte_hip.so gfx950 kernels + HIP backend (include/belewitte_hip.h)

Both are built in-tree by ``belewitte_amd.build`` into ``belewitte_amd/lib``.
Loading fails loudly when a library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class BelewitteError(RuntimeError):
    """A native call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class NativeLibraryMissing(ImportError):
    pass


# ------------------------------------------------------------------ structs

class BwLaunch(C.Structure):
    """bw_launch of include/belewitte_hip.h."""
    _fields_ = [("block", C.c_uint32), ("grid", C.c_uint32), ("warp", C.c_uint32),
                ("chunk", C.c_uint32), ("stream", vp)]


class BwGraphDev(C.Structure):
    """bw_graph_dev of include/belewitte_hip.h."""
    _fields_ = [("vertex_count", C.c_uint64), ("edge_count", C.c_uint64),
                ("offsets", vp), ("cols", vp), ("rev_offsets", vp), ("rev_cols", vp),
                ("el_in", vp), ("el_out", vp), ("rel_in", vp), ("rel_out", vp),
                ("sel", vp), ("rsel", vp), ("offset_bytes", C.c_int32), ("pad_", C.c_int32)]


class BwHipProps(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64),
                ("compute_units", C.c_int32), ("wave_size", C.c_int32),
                ("max_threads_per_block", C.c_int32),
                ("max_block_dims", C.c_int32 * 3), ("max_grid_dims", C.c_int32 * 3),
                ("shared_mem_per_block", C.c_uint64), ("total_global_mem", C.c_uint64),
                ("l2_cache_bytes", C.c_uint64), ("clock_khz", C.c_int32),
                ("memory_clock_khz", C.c_int32), ("memory_bus_width_bits", C.c_int32),
                ("pad_", C.c_int32)]


# --------------------------------------------------------------- signatures

_GRAPH_SIGS = {
    "bw_graph_last_error": (C.c_char_p, []),
    "bw_graph_rmat": (C.c_int, [C.POINTER(vp), C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                C.c_double, C.c_uint64, C.c_int]),
    "bw_graph_rmat_n": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                  C.c_double, C.c_uint64, C.c_int]),
    "bw_graph_erdos_renyi": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]),
    "bw_graph_grid2d": (C.c_int, [C.POINTER(vp), C.c_uint32, C.c_uint32, C.c_double, C.c_uint64]),
    "bw_graph_chain": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_int]),
    "bw_graph_star": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_int]),
    "bw_graph_mesh": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_uint64, C.c_int]),
    "bw_graph_from_edges": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_uint64, u32p, u32p, C.c_int, C.c_int]),
    "bw_graph_load": (C.c_int, [C.POINTER(vp), C.c_char_p]),
    "bw_graph_save": (C.c_int, [vp, C.c_char_p]),
    "bw_graph_free": (None, [vp]),
    "bw_graph_vertex_count": (C.c_uint64, [vp]),
    "bw_graph_edge_count": (C.c_uint64, [vp]),
    "bw_graph_is_undirected": (C.c_int, [vp]),
    "bw_graph_offsets": (u64p, [vp, C.c_int]),
    "bw_graph_cols": (u32p, [vp, C.c_int]),
    "bw_graph_offsets32": (C.c_int, [vp, C.c_int, u32p]),
    "bw_graph_max_out_degree_vertex": (C.c_uint32, [vp]),
    "bw_graph_degree_stats": (C.c_int, [vp, C.c_int, f64p]),
    "bw_graph_reorder_by_degree": (C.c_int, [C.POINTER(vp), vp, u32p]),
    "bw_hash64": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
}

_L = C.POINTER(BwLaunch)
_G = C.POINTER(BwGraphDev)

_HIP_SIGS = {
    "bw_hip_last_error": (C.c_char_p, []),
    "bw_hip_version": (C.c_char_p, []),
    "bw_hip_device_count": (C.c_int, [i32p]),
    "bw_hip_set_device": (C.c_int, [C.c_int32]),
    "bw_hip_get_device_props": (C.c_int, [C.c_int32, C.POINTER(BwHipProps)]),
    "bw_hip_dev_alloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "bw_hip_dev_free": (C.c_int, [vp]),
    "bw_hip_host_alloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "bw_hip_host_free": (C.c_int, [vp]),
    "bw_hip_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
    "bw_hip_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
    "bw_hip_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t]),
    "bw_hip_memcpy_async": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "bw_hip_fill_u32": (C.c_int, [vp, C.c_uint32, C.c_uint64, vp]),
    "bw_hip_device_sync": (C.c_int, []),
    "bw_hip_stream_create": (C.c_int, [C.POINTER(vp)]),
    "bw_hip_stream_destroy": (C.c_int, [vp]),
    "bw_hip_stream_sync": (C.c_int, [vp]),
    "bw_hip_event_create": (C.c_int, [C.POINTER(vp)]),
    "bw_hip_event_destroy": (C.c_int, [vp]),
    "bw_hip_event_record": (C.c_int, [vp, vp]),
    "bw_hip_event_sync": (C.c_int, [vp]),
    "bw_hip_event_elapsed_ms": (C.c_int, [f32p, vp, vp]),
    "bw_hip_compute_division": (C.c_int, [C.c_uint64, u64p, u64p]),
    # BFS
    "bw_bfs_edge_list": (C.c_int, [C.c_uint64, vp, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_rev_edge_list": (C.c_int, [C.c_uint64, vp, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_struct_edge_list": (C.c_int, [C.c_uint64, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_rev_struct_edge_list": (C.c_int, [C.c_uint64, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_vertex_push": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_vertex_pull": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_vertex_push_warp": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_vertex_pull_warp": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, C.c_int32, vp, C.c_int32, _L]),
    "bw_bfs_level": (C.c_int, [_G, C.c_int32, C.c_int32, vp, C.c_int32, vp, _L]),
    "bw_graph_relabel_by_degree": (C.c_int, [C.c_uint64, C.c_uint64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bw_build_edge_lists": (C.c_int, [C.c_uint64, C.c_uint64, vp, C.c_int32, vp, vp, vp, vp, vp]),
    "bw_bfs_run": (C.c_int, [_G, i32p, C.c_int32, C.c_int32, C.c_uint32, vp, u32p, f32p, C.c_int64, i64p, _L]),
    "bw_bfs_run_static": (C.c_int, [_G, i32p, C.c_int32, C.c_int32, C.c_uint32, vp, u32p, C.c_int64, i64p, _L]),
    # PageRank
    "bw_pr_zero_degrees": (C.c_int, [C.c_uint64, vp, _L]),
    "bw_pr_degrees_edge_list": (C.c_int, [C.c_uint64, vp, vp, _L]),
    "bw_pr_degrees_struct_edge_list": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, _L]),
    "bw_pr_degrees_csr": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, _L]),
    "bw_pr_degrees_rev_csr": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, _L]),
    "bw_pr_edge_list": (C.c_int, [C.c_uint64, vp, vp, vp, vp, vp, _L]),
    "bw_pr_rev_edge_list": (C.c_int, [C.c_uint64, vp, vp, vp, vp, vp, _L]),
    "bw_pr_struct_edge_list": (C.c_int, [C.c_uint64, vp, vp, vp, vp, _L]),
    "bw_pr_rev_struct_edge_list": (C.c_int, [C.c_uint64, vp, vp, vp, vp, _L]),
    "bw_pr_vertex_push": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, vp, _L]),
    "bw_pr_vertex_pull": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, vp, vp, _L]),
    "bw_pr_vertex_pull_nodiv": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, vp, _L]),
    "bw_pr_vertex_push_warp": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, vp, _L]),
    "bw_pr_vertex_pull_warp": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, vp, vp, _L]),
    "bw_pr_vertex_pull_warp_nodiv": (C.c_int, [C.c_uint64, vp, C.c_int32, vp, vp, vp, _L]),
    "bw_pr_consolidate": (C.c_int, [C.c_uint64, vp, vp, vp, C.c_int32, C.c_int32, C.c_int32, vp, vp, _L]),
    "bw_pr_dangling_mass": (C.c_int, [C.c_uint64, vp, vp, vp, _L]),
    "bw_pr_predivide": (C.c_int, [C.c_uint64, vp, vp, _L]),
    "bw_pr_run": (C.c_int, [_G, C.c_int32, C.c_int32, C.c_int32, vp, f64p, _L]),
    "bw_pr_pull_plan_create": (C.c_int, [C.POINTER(vp), C.c_uint64, vp, C.c_int32, vp]),
    "bw_pr_pull_plan_destroy": (C.c_int, [vp]),
    "bw_pr_pull_plan_info": (C.c_int, [vp, u32p]),
    "bw_pr_pull_iteration": (C.c_int, [vp, C.c_int32, vp, vp, vp, vp, _L]),
    "bw_pr_scatter_create": (C.c_int, [C.POINTER(vp), C.c_int32, C.c_uint64, C.c_uint64, vp, vp, C.c_int32, vp]),
    "bw_pr_scatter_destroy": (C.c_int, [vp]),
    "bw_pr_scatter_iteration": (C.c_int, [vp, vp, vp, vp, vp]),
    # fused pull PageRank
    "bw_pr_plan_create": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_uint64, vp, C.c_int32, vp]),
    "bw_pr_plan_destroy": (C.c_int, [vp]),
    "bw_pr_plan_info": (C.c_int, [vp, u32p, u32p, u32p]),
    "bw_pr_plan_degrees": (C.c_int, [vp, C.POINTER(vp)]),
    "bw_pr_plan_tune": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
    "bw_pr_plan_sort_windows": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "bw_pr_plan_bin": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
    "bw_pr_plan_bin_host": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
    "bw_pr_plan_bin_array": (C.c_int, [vp, C.c_int32, vp, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bw_pr_plan_bin_info": (C.c_int, [vp, u32p, u32p, u32p]),
    "bw_pr_fused_acc_len": (C.c_uint64, [C.c_int32]),
    "bw_pr_fused_run": (C.c_int, [_G, vp, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, f64p, f32p, _L]),
    # device-driven BFS engine
    "bw_bfs_engine_create": (C.c_int, [C.POINTER(vp), _G, vp, C.c_uint32]),
    "bw_bfs_engine_destroy": (C.c_int, [vp]),
    "bw_bfs_engine_levels": (C.c_int, [vp, C.POINTER(vp)]),
    "bw_bfs_engine_mlp_len": (C.c_int32, [C.c_int32]),
    "bw_bfs_engine_set_model": (C.c_int, [vp, f32p, C.c_int32]),
    "bw_bfs_engine_set_schedule": (C.c_int, [vp, i32p, C.c_int32]),
    "bw_bfs_engine_begin": (C.c_int, [vp, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, vp]),
    "bw_bfs_engine_level": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, vp]),
    "bw_bfs_engine_force": (C.c_int, [vp, C.c_int32, vp]),
    "bw_bfs_engine_status": (C.c_int, [vp, u32p, u32p, u32p, i32p, f32p, vp]),
    "bw_bfs_engine_snapshot": (C.c_int, [vp, C.c_int32, vp]),
    "bw_bfs_engine_step_forced": (C.c_int, [vp, C.c_int32, f32p, u32p, u32p, vp]),
    "bw_bfs_engine_run": (C.c_int, [vp, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    u32p, u64p, i32p, u32p, f32p, vp]),
    "bw_bfs_engine_set_tuning": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
    "bw_bfs_engine_set_queues": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
    "bw_bfs_engine_begin_multi": (C.c_int, [vp, u32p, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, vp]),
    "bw_bfs_engine_run_multi": (C.c_int, [vp, u32p, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, u32p, u64p, i32p, u32p, f32p, vp]),
    # multi-source BFS
    "bw_msbfs_create": (C.c_int, [C.POINTER(vp), _G, C.c_uint32]),
    "bw_msbfs_destroy": (C.c_int, [vp]),
    "bw_msbfs_set_tuning": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
    "bw_msbfs_run": (C.c_int, [vp, u32p, C.c_uint32, vp, C.c_int32, C.c_int32, u64p, u32p, f32p, vp]),
    # variant predictor
    "bw_mlp_params": (C.c_int32, [C.c_int32]),
    "bw_mlp_forward": (C.c_int, [vp, C.c_int32, vp, C.c_uint64, vp, vp]),
    "bw_mlp_grad": (C.c_int, [vp, C.c_int32, vp, vp, C.c_uint32, vp, vp, vp]),
    "bw_mlp_train": (C.c_int, [vp, C.c_int32, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                               C.c_uint32, vp, vp]),
}


def _load(name: str, sigs: dict, optional: tuple = ()) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run belewitte_amd.build.build_all() (or __graft_entry__.build())")
    lib = C.CDLL(str(path), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for fn, (res, args) in sigs.items():
        try:
            f = getattr(lib, fn)
        except AttributeError:
            if fn in optional:
                continue
            raise
        f.restype = res
        f.argtypes = args
    return lib


_graph_lib = None
_hip_lib = None


def graph_lib() -> C.CDLL:
    global _graph_lib
    if _graph_lib is None:
        _graph_lib = _load("libbwgraph.so", _GRAPH_SIGS)
    return _graph_lib


def hip_lib() -> C.CDLL:
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = _load("libbelewitte_hip.so", _HIP_SIGS)
    return _hip_lib


def hip_signatures() -> dict:
    return dict(_HIP_SIGS)


def graph_signatures() -> dict:
    return dict(_GRAPH_SIGS)


def check_graph(status: int) -> None:
    if status != 0:
        msg = graph_lib().bw_graph_last_error()
        raise BelewitteError(status, msg.decode() if msg else "unknown error")


def check_hip(status: int) -> None:
    if status != 0:
        msg = hip_lib().bw_hip_last_error()
        raise BelewitteError(status, msg.decode() if msg else "unknown error")


def launch(block: int = 0, grid: int = 0, warp: int = 0, chunk: int = 0, stream=None) -> BwLaunch:
    return BwLaunch(block, grid, warp, chunk, stream)
