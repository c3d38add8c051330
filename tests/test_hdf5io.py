"""HDF5 output and restart files in the reference's format (hdf5io_class.f03),
written by libqpio.so and read back with HDF5's own h5dump (an independent
reader), so the files are checked as a QuickPIC user's tools would see them:

* file names <dir><dataname>_<step %08d>.h5 (:393-395);
* root attributes NAME, TYPE, TIME, ITER, DT, TIME UNITS (wrattr_file
  :299-347), the AXIS group of a grid file, the dataset's UNITS and
  LONG_NAME (wrattr_dataset :349-369), the 100-character padding of the
  hdf5file members;
* Fortran dimension order: a (nx, ny, nz) Fortran array is a (nz, ny, nx)
  dataset with the same bytes;
* the raw beam datasets x1..x3 (position * delta + origin), p1..p3, q and the
  attribute tp (pwpart_3d_pipe :1018-1299);
* the restart dataset of Fortran dimensions (7, tp) and its round trip
  (wpart / rpart :1301-1421).
"""
import os
import re
import subprocess

import numpy as np
import pytest

qpdiag = pytest.importorskip("qpdiag")

H5DUMP = "/opt/conda/bin/h5dump"
needs_h5dump = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")


def dump(args):
    return subprocess.run([H5DUMP] + args, capture_output=True, text=True, check=True).stdout


def dump_data(path, dset, tmp_path):
    """The dataset's values through h5dump's binary export (little endian)."""
    out = str(tmp_path / "raw.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path], capture_output=True,
                   check=True)
    return np.fromfile(out, dtype="<f8")


def attr_block(text, name):
    """The text of attribute `name` up to the end of its DATA block."""
    m = re.search(r'ATTRIBUTE "%s" \{.*?DATA \{.*?\n\s*\}' % re.escape(name), text, re.S)
    assert m, name
    return m.group(0)


def test_library_builds_and_names_files():
    f = qpdiag.H5File(filename="./Fields/Ez/", dataname="ez", n=12)
    assert f.path() == "./Fields/Ez/ez_00000012.h5"
    g = qpdiag.H5File(filename="./RST/Beam-0001/", dataname="RST-beam0001-0000000000", n=3)
    assert g.path() == "./RST/Beam-0001/RST-beam0001-0000000000_00000003.h5"


@needs_h5dump
def test_grid_file_layout_and_attributes(tmp_path):
    f = qpdiag.H5File(filename=str(tmp_path) + "/Fields/Ez/", dataname="ez", rank=3,
                      units="mc\\omega_p/e", label="Electric Field", timeunits="1 / \\omega_p",
                      dt=10.0, n=3, t=30.0, axisname=("x", "y", "\\xi"),
                      axislabel=("x", "y", "\\xi"), axisunits=("c / \\omega_p",) * 3,
                      axismin=(-10.0, -9.0, 0.0), axismax=(10.0, 9.0, 10.0))
    nz, ny, nx = 5, 4, 3
    a = np.random.default_rng(1).standard_normal((nz, ny, nx))
    qpdiag.write_grid(f, a)
    p = f.path()
    assert os.path.exists(p)
    head = dump(["-H", "-A", p])
    # the dataset: Fortran (nx, ny, nz) appears as (nz, ny, nx)
    assert re.search(r'DATASET "ez" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE '
                     r'\{ \( 5, 4, 3 \) / \( 5, 4, 3 \) \}', head)
    assert np.array_equal(dump_data(p, "/ez", tmp_path), a.ravel())
    # root attributes
    for name in ("NAME", "TYPE", "TIME UNITS"):
        assert "STRSIZE 100;" in attr_block(head, name)
    assert '"ez' + " " * 98 + '"' in attr_block(head, "NAME")
    assert '"grid' + " " * 96 + '"' in attr_block(head, "TYPE")
    assert "(0): 30" in attr_block(head, "TIME") and "H5T_IEEE_F64LE" in attr_block(head, "TIME")
    assert "(0): 3" in attr_block(head, "ITER") and "H5T_STD_I32LE" in attr_block(head, "ITER")
    assert "(0): 10" in attr_block(head, "DT")
    # dataset attributes (hdf5file members, padded)
    assert '"mc\\omega_p/e' in head and '"Electric Field' + " " * 86 + '"' in head
    # AXIS group: [min, max] per axis with trimmed names
    for i, (lo, hi, nm) in enumerate(((-10, 10, "x"), (-9, 9, "y"), (0, 10, "\\xi")), start=1):
        v = dump_data(p, f"/AXIS/AXIS{i}", tmp_path)
        assert np.array_equal(v, [lo, hi])
        blk = dump(["-A", "-d", f"/AXIS/AXIS{i}", p])
        assert f'"{nm}"' in blk and '"linear"' in blk and '"c / \\omega_p"' in blk


@needs_h5dump
def test_slice_file_is_rank_2(tmp_path):
    f = qpdiag.H5File(filename=str(tmp_path) + "/Fields/Ez_slice0001/", dataname="ezslicexz",
                      rank=2, n=0, t=0.0)
    a = np.arange(12.0).reshape(4, 3)          # (nz, nx): Fortran (nx, nz)
    qpdiag.write_grid(f, a)
    head = dump(["-H", f.path()])
    assert "SIMPLE { ( 4, 3 ) / ( 4, 3 ) }" in head
    assert 'DATASET "AXIS2"' in head and 'DATASET "AXIS3"' not in head
    assert np.array_equal(dump_data(f.path(), "/ezslicexz", tmp_path), a.ravel())


@needs_h5dump
def test_raw_beam_file(tmp_path):
    rng = np.random.default_rng(2)
    part = rng.uniform(1.0, 30.0, (101, 7))
    f = qpdiag.H5File(filename=str(tmp_path) + "/Beam0001/Raw/", dataname="raw", ty="particles",
                      units="", label="Beam Raw", n=1, t=10.0)
    delta, origin = (0.5, 0.25, 0.125), (-10.0, -8.0, 0.0)
    qpdiag.write_raw(f, part, 3, delta, origin)
    p = f.path()
    txt = dump(["-A", p])
    assert "(0): 33" in attr_block(txt, "tp")            # int(101 / 3) records
    assert 'GROUP "AXIS"' not in txt                      # particles: no AXIS group
    sel = part[0:33 * 3:3]
    for i in range(3):
        got = dump_data(p, f"/x{i + 1}", tmp_path)
        assert np.array_equal(got, sel[:, i] * delta[i] + origin[i])
        assert np.array_equal(dump_data(p, f"/p{i + 1}", tmp_path), sel[:, 3 + i])
    assert np.array_equal(dump_data(p, "/q", tmp_path), sel[:, 6])
    h = dump(["-H", "-p", "-d", "/x1", p])
    assert "H5S_UNLIMITED" in h and "CHUNKED ( 33 )" in h


@needs_h5dump
def test_restart_round_trip(tmp_path):
    part = np.random.default_rng(3).standard_normal((57, 7))
    f = qpdiag.H5File(filename=str(tmp_path) + "/RST/Beam-0001/",
                      dataname="RST-beam0001-0000000000", ty="rst", n=4, t=40.0)
    qpdiag.write_rst(f, part)
    p = f.path()
    head = dump(["-H", "-A", p])
    assert "SIMPLE { ( 57, 7 ) / ( 57, 7 ) }" in head      # Fortran (7, tp)
    assert "(0): 57" in attr_block(dump(["-A", p]), "tp")
    assert np.array_equal(dump_data(p, "/RST-beam0001-0000000000", tmp_path), part.ravel())
    back = qpdiag.read_rst(f, 100)
    assert np.array_equal(back, part)
    with pytest.raises(RuntimeError):
        qpdiag.read_rst(f, 10)                             # too many records for the room


def test_empty_restart(tmp_path):
    f = qpdiag.H5File(filename=str(tmp_path) + "/", dataname="RST", ty="rst", n=1)
    qpdiag.write_rst(f, np.zeros((0, 7)))
    assert qpdiag.read_rst(f, 5).shape == (0, 7)
