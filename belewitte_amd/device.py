"""Device memory and device-resident graph representations.

:class:`DeviceArray` is one hipMalloc'ed buffer owned through the C ABI.
:class:`DeviceGraph` uploads a :class:`~belewitte_amd.graph.Graph` and
builds the representations of GraphLoader.hpp (CSR, reverse CSR, EdgeList,
StructEdgeList, each direction) lazily: only what a variant asks for is
materialised, and the edge-list forms are expanded on the GPU from the CSR
instead of on the host (GraphLoader.hpp:115-156 builds every form on the
host for both directions).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def device_count() -> int:
    n = C.c_int32(0)
    N.check_hip(N.hip_lib().bw_hip_device_count(C.byref(n)))
    return int(n.value)


def require_gpu() -> None:
    if device_count() < 1:
        raise N.BelewitteError(3, "no GPU visible: the HIP kernels need a device")


def set_device(device: int = 0) -> None:
    N.check_hip(N.hip_lib().bw_hip_set_device(device))


def device_props(device: int = 0) -> dict:
    p = N.BwHipProps()
    N.check_hip(N.hip_lib().bw_hip_get_device_props(device, C.byref(p)))
    return {
        "name": p.name.decode(), "arch": p.arch.decode(), "compute_units": p.compute_units,
        "wave_size": p.wave_size, "max_threads_per_block": p.max_threads_per_block,
        "shared_mem_per_block": p.shared_mem_per_block, "total_global_mem": p.total_global_mem,
        "l2_cache_bytes": p.l2_cache_bytes, "clock_khz": p.clock_khz,
        "memory_clock_khz": p.memory_clock_khz, "memory_bus_width_bits": p.memory_bus_width_bits,
    }


def synchronize() -> None:
    N.check_hip(N.hip_lib().bw_hip_device_sync())


class DeviceArray:
    """A typed device buffer."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        N.check_hip(N.hip_lib().bw_hip_dev_alloc(C.byref(p), self.nbytes))
        self._ptr = p

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(arr)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.size != self.count:
            raise ValueError(f"size mismatch {a.size} != {self.count}")
        N.check_hip(N.hip_lib().bw_hip_memcpy_h2d(self._ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def download(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.count, dtype=self.dtype)
        N.check_hip(N.hip_lib().bw_hip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes))
        return out

    def fill_u32(self, value: int, stream=None) -> None:
        if self.dtype.itemsize != 4:
            raise TypeError("fill_u32 needs a 32-bit element type")
        N.check_hip(N.hip_lib().bw_hip_fill_u32(self._ptr, int(value) & 0xFFFFFFFF, self.count, stream))

    def free(self) -> None:
        if self._ptr and self._ptr.value:
            N.check_hip(N.hip_lib().bw_hip_dev_free(self._ptr))
            self._ptr = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        N.check_hip(N.hip_lib().bw_hip_event_create(C.byref(p)))
        self._p = p

    def record(self, stream=None):
        N.check_hip(N.hip_lib().bw_hip_event_record(self._p, stream))

    def synchronize(self):
        N.check_hip(N.hip_lib().bw_hip_event_sync(self._p))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        N.check_hip(N.hip_lib().bw_hip_event_elapsed_ms(C.byref(ms), self._p, end._p))
        return float(ms.value)

    def __del__(self):
        try:
            N.hip_lib().bw_hip_event_destroy(self._p)
        except Exception:
            pass


REPRESENTATIONS = ("csr", "rev_csr", "edge_list", "rev_edge_list", "struct_edge_list",
                   "rev_struct_edge_list")


class DeviceGraph:
    """Device-resident representations of one host graph."""

    def __init__(self, g, offset_bytes: int = 4, reps=("csr", "rev_csr")):
        if offset_bytes not in (4, 8):
            raise ValueError("offset_bytes must be 4 or 8")
        if offset_bytes == 4 and g.edge_count > 0xFFFFFFFF:
            raise ValueError("edge count needs 64-bit offsets")
        self.host = g
        self.vertex_count = g.vertex_count
        self.edge_count = g.edge_count
        self.offset_bytes = offset_bytes
        self.undirected = g.undirected
        self._a: dict[str, DeviceArray] = {}
        self.desc = N.BwGraphDev()
        self.desc.vertex_count = self.vertex_count
        self.desc.edge_count = self.edge_count
        self.desc.offset_bytes = offset_bytes
        for r in reps:
            self.ensure(r)

    def _offsets(self, reverse: bool) -> np.ndarray:
        if self.offset_bytes == 4:
            return self.host.offsets32(reverse)
        return np.ascontiguousarray(self.host.offsets(reverse))

    def ensure(self, rep: str) -> None:
        """Materialises one representation (GraphLoader::loadGraph for it)."""
        if rep not in REPRESENTATIONS:
            raise ValueError(f"unknown representation {rep!r}")
        d = self.desc
        if rep in ("csr", "rev_csr"):
            rev = rep == "rev_csr"
            key_o, key_c = ("roff", "rcol") if rev else ("off", "col")
            if key_o in self._a:
                return
            if rev and self.undirected and "off" in self._a:
                self._a["roff"], self._a["rcol"] = self._a["off"], self._a["col"]
            else:
                self._a[key_o] = DeviceArray.from_host(self._offsets(rev))
                self._a[key_c] = DeviceArray.from_host(self.host.cols(rev)) if self.edge_count \
                    else DeviceArray(1, np.uint32)
            if rev:
                d.rev_offsets, d.rev_cols = self._a["roff"].ptr, self._a["rcol"].ptr
            else:
                d.offsets, d.cols = self._a["off"].ptr, self._a["col"].ptr
            return
        rev = rep.startswith("rev_")
        self.ensure("rev_csr" if rev else "csr")
        off = self._a["roff" if rev else "off"]
        col = self._a["rcol" if rev else "col"]
        E = max(self.edge_count, 1)
        lib = N.hip_lib()
        if rep.endswith("struct_edge_list"):
            key = "rsel" if rev else "sel"
            if key in self._a:
                return
            pairs = DeviceArray(2 * E, np.uint32)
            N.check_hip(lib.bw_build_edge_lists(self.vertex_count, self.edge_count, off.ptr, self.offset_bytes,
                                                 col.ptr, None, None, pairs.ptr, None))
            self._a[key] = pairs
            setattr(d, key, pairs.ptr)
        else:
            ki, ko = ("rel_in", "rel_out") if rev else ("el_in", "el_out")
            if ki in self._a:
                return
            a_in, a_out = DeviceArray(E, np.uint32), DeviceArray(E, np.uint32)
            N.check_hip(lib.bw_build_edge_lists(self.vertex_count, self.edge_count, off.ptr, self.offset_bytes,
                                                col.ptr, a_in.ptr, a_out.ptr, None, None))
            self._a[ki], self._a[ko] = a_in, a_out
            setattr(d, ki, a_in.ptr)
            setattr(d, ko, a_out.ptr)
        synchronize()

    def relabel_by_degree(self):
        """The graph relabelled on the device by decreasing out-degree, ties
        by id (bw_graph_relabel_by_degree): the same graph as
        Graph.reorder_by_degree() gives on the host, built from the arrays
        already here with no copy to the host.  Returns (DeviceGraph, perm)
        with perm (host uint32[V]) the new id of each old vertex."""
        if self.offset_bytes != 4:
            raise ValueError("the device relabel reads uint32 offsets")
        self.ensure("csr")
        self.ensure("rev_csr")
        V, E = self.vertex_count, self.edge_count
        out = DeviceGraph.__new__(DeviceGraph)
        out.host = None
        out.vertex_count, out.edge_count, out.offset_bytes, out.undirected = V, E, 4, self.undirected
        out._a = {}
        out.desc = N.BwGraphDev()
        out.desc.vertex_count, out.desc.edge_count, out.desc.offset_bytes = V, E, 4
        out._a["off"] = DeviceArray(V + 1, np.uint32)
        out._a["col"] = DeviceArray(max(E, 1), np.uint32)
        if self.undirected:
            out._a["roff"], out._a["rcol"] = out._a["off"], out._a["col"]
        else:
            out._a["roff"] = DeviceArray(V + 1, np.uint32)
            out._a["rcol"] = DeviceArray(max(E, 1), np.uint32)
        perm = DeviceArray(max(V, 1), np.uint32)
        rev = (None, None, None, None) if self.undirected else (
            self._a["roff"].ptr, self._a["rcol"].ptr, out._a["roff"].ptr, out._a["rcol"].ptr)
        try:
            N.check_hip(N.hip_lib().bw_graph_relabel_by_degree(
                V, E, self._a["off"].ptr, self._a["col"].ptr, rev[0], rev[1], out._a["off"].ptr,
                out._a["col"].ptr, rev[2], rev[3], perm.ptr, None))
            p = perm.download()[:V]
        except Exception:
            out.free()
            raise
        finally:
            perm.free()
        out.desc.offsets, out.desc.cols = out._a["off"].ptr, out._a["col"].ptr
        out.desc.rev_offsets, out.desc.rev_cols = out._a["roff"].ptr, out._a["rcol"].ptr
        return out, p

    _KEY = {"csr": "off", "rev_csr": "roff", "edge_list": "el_in", "rev_edge_list": "rel_in",
            "struct_edge_list": "sel", "rev_struct_edge_list": "rsel"}

    def has(self, rep: str) -> bool:
        """Whether a representation is resident."""
        return self._KEY[rep] in self._a

    def array(self, key: str) -> DeviceArray:
        return self._a[key]

    @property
    def ref(self):
        return C.byref(self.desc)

    def free(self) -> None:
        seen = set()
        for a in self._a.values():
            if id(a) not in seen:
                seen.add(id(a))
                a.free()
        self._a.clear()
