"""The gather-rate probe behind the L2 request ceiling of the PageRank gathers."""
import ctypes as C

import numpy as np
import pytest

from belewitte_amd import profiling


def test_probe_library_is_separate_from_the_kernel_abi():
    """The probes are exported by libbw_probes.so only, not by the kernel
    library whose header lists the path's entry points."""
    from belewitte_amd import _native as N
    lib = profiling.probes_lib()
    assert hasattr(lib, "bw_probe_copy") and hasattr(lib, "bw_probe_gather")
    for name in ("bw_hip_copy_kernel", "bw_hip_gather_probe", "bw_probe_copy"):
        assert not hasattr(N.hip_lib(), name)


def test_indices_share_lines():
    idx = profiling.gather_indices(1 << 12, 1 << 16, 4, seed=3)
    assert idx.max() < (1 << 16)
    lines = idx.astype(np.int64) // profiling.LINE_FLOATS
    # groups of 4 consecutive entries read one line, each its own word
    assert np.all(lines.reshape(-1, 4) == lines.reshape(-1, 4)[:, :1])
    assert np.array_equal(idx % profiling.LINE_FLOATS, np.arange(idx.size) % profiling.LINE_FLOATS)
    with pytest.raises(ValueError):
        profiling.gather_indices(64, 1 << 10, 3)


@pytest.mark.gpu
def test_probe_sums_match_numpy(gpu):
    from belewitte_amd.device import DeviceArray
    threads, n, tlen = 256 * 4, 256 * 4 * 16 * 3, 4096
    table = (np.arange(tlen) % 7).astype(np.float32)
    idx = profiling.gather_indices(n, tlen, 8, seed=5)
    idx[::97] = tlen + 5  # out of range: reads 0
    d_t, d_i, d_o = DeviceArray(tlen, np.float32), DeviceArray(n, np.uint32), DeviceArray(threads, np.float32)
    d_t.upload(table)
    d_i.upload(idx)
    ms = C.c_float()
    lib = profiling.probes_lib()
    profiling.check_probe(lib.bw_probe_gather(d_i.ptr, n, d_t.ptr, tlen, d_o.ptr, threads, C.byref(ms), None))
    got = d_o.download()[:threads]
    vals = np.where(idx < tlen, table[np.minimum(idx, tlen - 1)], 0.0).astype(np.float64)
    want = vals.reshape(-1, threads).sum(axis=0)  # entry r * threads + t belongs to thread t
    assert np.array_equal(got.astype(np.float64), want)
    assert ms.value > 0
    # bad shapes are refused before launch
    assert lib.bw_probe_gather(d_i.ptr, n + 1, d_t.ptr, tlen, d_o.ptr, threads, None, None) != 0
    assert lib.bw_probe_gather(d_i.ptr, n, d_t.ptr, tlen, d_o.ptr, 100, None, None) != 0
    for a in (d_t, d_i, d_o):
        a.free()
