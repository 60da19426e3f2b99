"""The checkpoint container: the package's own HDF5 writer and reader.

Pins:
- the reader against a file libhdf5 wrote: scipy ships a MATLAB v7.3 file (HDF5 after a 512-byte
  user block) holding linspace(0, 2 pi, 9) with the attribute MATLAB_class = "double"; the same
  MATLAB session's v7.1 file, decoded by scipy's own reader, gives the expected values;
- the writer's encodings byte for byte against the messages libhdf5 wrote in that file (the
  binary64 datatype, the dataspace, the attribute message up to its character-set bit) and its
  superblock's format fields;
- round trips of every value type the checkpoint stores, groups spread over several
  symbol-table nodes, and refusals of what the reader does not take.
"""
import math
import struct
from pathlib import Path

import numpy as np
import pytest

MAT_DIR = Path(np.__file__).resolve().parent.parent / "scipy" / "io" / "matlab" / "tests" / "data"
MAT_HDF5 = MAT_DIR / "testhdf5_7.4_GLNX86.mat"
MAT_V5 = MAT_DIR / "testdouble_7.1_GLNX86.mat"


@pytest.fixture(scope="module")
def h5(g):
    from gbg_amd import hdf5
    return hdf5


def _libhdf5_file():
    if not MAT_HDF5.exists():
        pytest.skip("scipy's MATLAB test data is not installed")
    return MAT_HDF5


def test_reader_reads_a_file_libhdf5_wrote(h5):
    import scipy.io
    path = _libhdf5_file()
    assert h5.is_hdf5(path)
    root = h5.read_file(path)
    assert list(root.members) == ["testdouble"]
    ds = root.members["testdouble"]
    assert ds.attrs == {"MATLAB_class": "double"}
    expected = scipy.io.loadmat(MAT_V5)["testdouble"]          # (1, 9) in MATLAB's column-major view
    assert ds.data.shape == (9, 1) and ds.data.dtype == np.float64
    assert np.array_equal(ds.data.ravel(), expected.ravel())
    assert np.array_equal(ds.data.ravel(), np.linspace(0, 2 * np.pi, 9))


def test_writer_encodings_match_the_bytes_libhdf5_wrote(h5):
    path = _libhdf5_file()
    raw = path.read_bytes()
    sb = raw[512:512 + 96]
    with open(path, "rb") as f:
        r = h5._Reader(f)
        entries = r._symbol_table(*struct.unpack_from("<QQ", sb, 80))
        assert entries[0][0] == "testdouble"
        msgs = r.messages(entries[0][1])
    by_type = {}
    for t, _, body in msgs:
        by_type.setdefault(t, body)
    # the binary64 datatype and the rank-2 dataspace, byte for byte
    dt = h5._dtype_message(np.dtype("<f8"))
    assert by_type[h5._DATATYPE][:len(dt)] == dt
    ds = h5._dataspace_message((9, 1))
    assert by_type[h5._DATASPACE][:len(ds)] == ds
    # the attribute message: libhdf5 wrote an ASCII string, this writer marks it UTF-8 (as HDF5.jl does)
    ours = h5._attribute_message("MATLAB_class", "double")[8:]
    theirs = bytearray(by_type[h5._ATTRIBUTE][:len(ours)])
    assert len(ours) == len(by_type[h5._ATTRIBUTE])
    theirs[24 + 1] |= 0x10
    assert bytes(theirs) == ours
    # superblock version 0: free-space, root-entry and shared-header versions, 8-byte offsets and
    # lengths, group leaf and internal node K, and the root entry's cached symbol table
    w = _written(h5, tmp=None)
    assert w[8:20] == sb[8:20]
    assert struct.unpack_from("<I", w, 72)[0] == 1 == struct.unpack_from("<I", sb, 72)[0]


def _written(h5, tmp=None, root=None):
    import tempfile
    if root is None:
        root = h5.Group()
        root.members["testdouble"] = h5.Dataset(np.linspace(0, 2 * np.pi, 9).reshape(9, 1),
                                                {"MATLAB_class": "double"})
    with tempfile.TemporaryDirectory() as d:
        p = Path(d) / "f.h5"
        h5.write_file(p, root)
        return p.read_bytes()


def test_the_writer_lays_out_libhdf5_structures(h5, tmp_path):
    """Symbol-table nodes and B-tree nodes are allocated at libhdf5's full sizes for leaf K 4 and
    node K 16, the heap starts with the empty name and ends its free list with offset 1."""
    root = h5.Group()
    root.members["x"] = np.arange(3.0)
    p = tmp_path / "s.h5"
    h5.write_file(p, root)
    b = p.read_bytes()
    heap = b.index(b"HEAP")
    size, free, data = struct.unpack_from("<QQQ", b, heap + 8)
    assert free == 1 and data == heap + 32 and b[data:data + 8] == b"\0" * 8 and size == 16
    assert b[data + 8:data + 10] == b"x\0"
    snod, tree = b.index(b"SNOD"), b.index(b"TREE")
    assert tree - snod == h5._SNOD_SIZE == 328
    assert struct.unpack_from("<QQ", b, 80) == (tree, heap)
    assert len(b) == struct.unpack_from("<Q", b, 40)[0]         # the end-of-file address


def test_round_trip_of_every_checkpoint_value(h5, tmp_path):
    root = h5.Group()
    root.attrs.update(title="DNS", n=512, side=2 * math.pi, top=math.inf, flag=True, off=False,
                      times=[0.0, 1e-3, 2e-3], schedule=[1, 0, 3], empty=[], unicode="ε = 0.1")
    sim = h5.Group()
    root.members["simulation"] = sim
    rng = np.random.default_rng(5)
    arrays = {
        "c128": rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5)),
        "c64": (rng.standard_normal((2, 7)) + 1j * rng.standard_normal((2, 7))).astype(np.complex64),
        "f32": rng.standard_normal((5, 3)).astype(np.float32),
        "f64": rng.standard_normal(11),
        "i32": np.arange(-4, 4, dtype=np.int32),
        "i64": np.arange(6, dtype=np.int64).reshape(2, 3),
        "zero": np.zeros((0, 3)),
        "scalar": np.array(2.5),
    }
    for k, v in arrays.items():
        sim.members[k] = v
    sim.members["empty_group"] = h5.Group()
    sim.members["empty_group"].attrs["type"] = "GBGHomogeneousTurbulence.NoForcing"
    many = h5.Group()                                            # 40 members: five symbol-table nodes
    for i in range(40):
        many.members[f"sample{i + 1}"] = np.full((3, 2), float(i))
    sim.members["trail"] = many
    p = tmp_path / "r.h5"
    h5.write_file(p, root)
    back = h5.read_file(p)
    a = back.attrs
    assert a["title"] == "DNS" and a["n"] == 512 and isinstance(a["n"], int)
    assert a["side"] == 2 * math.pi and a["top"] == math.inf
    assert a["flag"] == 1 and a["off"] == 0
    assert np.array_equal(a["times"], [0.0, 1e-3, 2e-3]) and a["schedule"].tolist() == [1, 0, 3]
    assert a["empty"].shape == (0,) and a["unicode"] == "ε = 0.1"
    s = back.members["simulation"]
    for k, v in arrays.items():
        got = s.members[k].data
        assert got.dtype == v.dtype and got.shape == v.shape and np.array_equal(got, v), k
    assert s.members["empty_group"].members == {}
    assert s.members["empty_group"].attrs["type"] == "GBGHomogeneousTurbulence.NoForcing"
    t = s.members["trail"].members
    assert sorted(t) == sorted(f"sample{i + 1}" for i in range(40))
    assert all(np.array_equal(t[f"sample{i + 1}"].data, np.full((3, 2), float(i))) for i in range(40))


def test_refusals(h5, tmp_path):
    p = tmp_path / "not.h5"
    p.write_bytes(b"PK\x03\x04" + b"\0" * 600)
    assert not h5.is_hdf5(p)
    with pytest.raises(h5.FormatError, match="no HDF5 signature"):
        h5.read_file(p)
    root = h5.Group()
    root.members["a/b"] = np.zeros(2)
    with pytest.raises(ValueError, match="not a valid link name"):
        h5.write_file(tmp_path / "bad.h5", root)
    root = h5.Group()
    root.members["s"] = np.array(["text"])
    with pytest.raises(ValueError, match="cannot store a dataset"):
        h5.write_file(tmp_path / "bad.h5", root)
    root = h5.Group()
    root.members.update({f"m{i}": np.zeros(1) for i in range(257)})
    with pytest.raises(ValueError, match="at most 256 members"):
        h5.write_file(tmp_path / "bad.h5", root)
    root = h5.Group()
    root.attrs["big"] = list(range(9000))
    with pytest.raises(ValueError, match="larger than 64 KiB"):
        h5.write_file(tmp_path / "bad.h5", root)


def test_the_domain_keeps_its_dealiasing_through_the_file(g, tmp_path):
    """write_domain / read_domain (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:119-154): the tag, HouLi's
    order, its strength only when set and its materialized flag only when true."""
    from gbg_amd import checkpoint as ck
    cases = [(g.TwoThirds(), np.float64), (g.PhaseShift(), np.float32),
             (g.HouLi(order=36, strength=63.463), np.float64), (g.HouLi(order=24), np.float32),
             (g.HouLi(order=36, materialized=True), np.float64)]
    s = ck.Store()
    for i, (d, et) in enumerate(cases):
        ck.write_domain(s.root, f"d{i}", g.FixedDomain(grid_points_per_axis=32, side_length=1.5, dealiasing=d,
                                                        element_type=et))
    s.save(tmp_path / "d.h5")
    r = ck.Store.load(tmp_path / "d.h5")
    assert "dealiasing_strength" not in r.root.group("d3", "t").attrs
    assert "dealiasing_materialized" not in r.root.group("d2", "t").attrs
    for i, (d, et) in enumerate(cases):
        dom = ck.read_domain(r.root, f"d{i}")
        assert repr(dom.dealiasing) == repr(d) and dom.element_type is et and dom.side_length == 1.5
        if isinstance(d, g.HouLi):
            assert dom.dealiasing.materialized == d.materialized
    bad = ck.Store()
    ck.write_domain(bad.root, "d", g.FixedDomain(grid_points_per_axis=32, side_length=1.0,
                                                  dealiasing=g.TwoThirds(), element_type=np.float64))
    bad.root.group("d", "t").attrs["dealiasing"] = "Spectral"
    with pytest.raises(ValueError, match="unrecognized dealiasing tag 'Spectral'"):
        ck.read_domain(bad.root, "d")
