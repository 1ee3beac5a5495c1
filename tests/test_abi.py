"""The C-ABI libraries load and export every declared symbol; host-side input checks
of gst_open run before any device call (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gpu4gst_1_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^[\w\s\*]*?\b(gst[d]?_\w+)\s*\(", src, re.M)))


def test_libgst_exports_all_declared_symbols():
    lib = ctypes.CDLL(engine.lib_path())
    names = _declared("gst.h")
    assert "gst_open" in names and "gst_solve" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(engine.EXPORTS) <= set(names)


def test_libgstdata_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "gpu4gst-1_amd", "libgstdata.so"))
    for name in _declared("gst_data.h"):
        assert hasattr(lib, name), name


def test_liboracle_exports():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    for name in _declared_oracle():
        assert hasattr(lib, name), name


def _declared_oracle():
    src = open(os.path.join(ROOT, "oracle", "oracle.h")).read()
    return sorted(set(re.findall(r"^int\s+(orc_\w+)\s*\(", src, re.M)))


def test_arena_bytes():
    assert engine.arena_bytes(1000, 5) == 1000 * 32 * 4
    assert engine.arena_bytes(4_800_000, 7, 5) == 4_800_000 * 128 * 6 * 4  # 14.7 GB
    assert engine.arena_bytes(3_070_000, 5) == 3_070_000 * 32 * 4


def test_strerror():
    L = engine.lib()
    for code in (0, 1, 2, 3, -1, -2, -3, -4, -5, -6):
        assert L.gst_strerror(code)


def _open(beg, adj, w, offs, verts):
    L = engine.lib()
    beg = np.asarray(beg, np.int64)
    adj = np.asarray(adj, np.int32)
    w = np.asarray(w, np.int32)
    offs = np.asarray(offs, np.int64)
    verts = np.asarray(verts, np.int32)
    csr = engine.GstCsr(len(beg) - 1, len(adj), beg.ctypes.data, adj.ctypes.data, w.ctypes.data)
    grp = engine.GstGroups(len(offs) - 1, offs.ctypes.data, verts.ctypes.data)
    h = ctypes.c_void_p()
    rc = L.gst_open(ctypes.byref(csr), ctypes.byref(grp), ctypes.byref(engine.GstOpts()), ctypes.byref(h))
    if rc == 0:
        L.gst_close(h)
    return rc


def test_open_rejects_malformed_input():
    ok = dict(beg=[0, 1, 2], adj=[1, 0], w=[5, 5], offs=[0, 1], verts=[0])
    bad_weight = dict(ok, w=[0, 0])
    bad_adj = dict(ok, adj=[7, 0])
    bad_beg = dict(ok, beg=[0, 2, 1])
    bad_group = dict(ok, verts=[9])
    unsorted = dict(ok, offs=[0, 2], verts=[1, 0])
    for case in (bad_weight, bad_adj, bad_beg, bad_group, unsorted):
        assert _open(**case) == engine.GST_BAD_INPUT, case


def test_open_rejects_arc_count_past_limit():
    """Arc offsets are 32-bit on the device with one chunk of room: m <= 2^31 - 2^25. The
    size check comes before any array is read, so small arrays stand in for the CSR."""
    L = engine.lib()
    beg = np.asarray([0, 1, 2], np.int64)
    adj = np.asarray([1, 0], np.int32)
    w = np.asarray([5, 5], np.int32)
    offs = np.asarray([0, 1], np.int64)
    verts = np.asarray([0], np.int32)
    grp = engine.GstGroups(1, offs.ctypes.data, verts.ctypes.data)
    limit = (1 << 31) - (1 << 25)
    for m, msg in ((limit + 1, b"bad CSR sizes"), (limit, b"beg_pos must run from 0 to m")):
        csr = engine.GstCsr(2, m, beg.ctypes.data, adj.ctypes.data, w.ctypes.data)
        h = ctypes.c_void_p()
        rc = L.gst_open(ctypes.byref(csr), ctypes.byref(grp), ctypes.byref(engine.GstOpts()), ctypes.byref(h))
        assert rc == engine.GST_BAD_INPUT, (m, rc)
        assert L.gst_last_error().startswith(msg), (m, L.gst_last_error())


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_open_without_device():
    assert _open([0, 1, 2], [1, 0], [5, 5], [0, 1], [0]) == engine.GST_NO_DEVICE


@pytest.mark.parametrize("ctype, cname", [("GstResult", "gst_result"), ("GstOpts", "gst_opts"),
                                          ("GstCsr", "gst_csr"), ("GstGroups", "gst_groups"),
                                          ("GstQuery", "gst_query"), ("GstInfo", "gst_info")])
def test_ctypes_mirror_matches_header_layout(tmp_path, ctype, cname):
    """Every field of the ctypes mirror sits at the offset the C compiler gives it."""
    cls = getattr(engine, ctype)
    fields = [f for f, _ in cls._fields_]
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gst.h"', 'int main(void) {',
             f'  printf("size %zu\\n", sizeof({cname}));']
    lines += [f'  printf("{f} %zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                       text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(cls)
    for f in fields:
        assert int(out[f]) == getattr(cls, f).offset, f
