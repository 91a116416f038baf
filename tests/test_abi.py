"""The C ABI library loads and exports every symbol include/smgpu.h
declares (CPU; no GPU compute)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "smgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(sm_[a-z0-9_]+)\s*\(", src))
    return sorted(n for n in names if not n.endswith("_fn"))


def test_library_exports_header():
    from seismicmesh_amd import _lib
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(_lib.lib, n), n
        assert n in _lib.EXPORTED, n


def test_device_query_and_errors():
    from seismicmesh_amd import _lib
    n = _lib.device_count()
    assert n >= 0
    assert _lib.lib.sm_version() >= 100
    # a bad argument is reported, not crashed on
    rc = _lib.lib.sm_unique_edges(None, -1, 0, None, None, None, 0, None)
    assert rc < 0
    assert "negative" in _lib.last_error()


def test_tape_lowering():
    from seismicmesh_amd import geometry as geo
    from seismicmesh_amd import _lib
    u = geo.Union([geo.Disk([0, 0], 1), geo.Disk([1, 0], 1), geo.Disk([2, 0], 1)])
    ops = [n.op for n in u.tape()]
    assert ops == [_lib.OP_DISK, _lib.OP_DISK, _lib.OP_UNION, _lib.OP_DISK, _lib.OP_UNION]
    many = geo.Union([geo.Ball([i, 0, 0], 0.5) for i in range(60)])
    geo.validate_tape(many.tape())  # left folds keep the stack at depth 2
    d = geo.Difference([geo.Disk([0, 0], 1), geo.Disk([1, 0], 1)], smoothness=0.1)
    assert [n.op for n in d.tape()][-1] == _lib.OP_SDIFF_STEP
    r = geo.Rectangle((0.0, 1.0, 0.0, 2.0), rotate=[0.3, 0, 0], translate=[1.0, 2.0])
    nd = r.tape()[0]
    assert nd.flags == _lib.XF_TRANSLATE | _lib.XF_ROT_X
