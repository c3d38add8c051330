"""Diagnostic output and beam restart files of a deck run, in QuickPIC's
HDF5 format, through libqpio.so (include/qpio.h).

Mirrors the reference's diagnostics set-up and dump (init_diag,
simulation_class.f03:771-1283; diag_simulation, :1285-1352):

* field.diag names ex, ey, ez (bexyz), bx, by, bz (bbxyz) and psi (the psi
  slices, kept by the slice loop when asked for) -> ./Fields/<Name>/<name>_
  <step>.h5, or 2-D slices ./Fields/<Name>_slice<k>/<name>slice<plane>_...;
* species(i).diag names charge (the smoothed species charge of cbq) and
  jx, jy, jz (the current of the last corrector pass) ->
  ./Species<i>/<Name>/...;
* beam(i).diag names charge (the 3-D beam charge) and raw (every sample-th
  particle) -> ./Beam<i>/Charge/... and ./Beam<i>/Raw/raw_...;
* simulation.dump_restart: the beam particles every ndump_restart steps ->
  ./RST/Beam-<i>/RST-beam<i>-0000000000_<step>.h5, which a run with
  read_restart reads back (rpart) before its first step.

Slices are chosen as in the deck: [plane, position] with plane 'xy' (a xi
plane at position 1..nz), 'xz' (y node position-1) or 'yz' (x node
position-1).  Data are downloaded from the GPU at the step's end, after the
beam push (the reference's order) and before the plasma renewal.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

import qpslice as qps

_HERE = os.path.dirname(os.path.abspath(__file__))
IOPATH = os.path.join(_HERE, "libqpio.so")
_io = None


class QpioFile(C.Structure):
    _fields_ = [("filename", C.c_char_p), ("dataname", C.c_char_p), ("timeunits", C.c_char_p),
                ("ty", C.c_char_p), ("units", C.c_char_p), ("label", C.c_char_p),
                ("n", C.c_int), ("rank", C.c_int), ("t", C.c_double), ("dt", C.c_double),
                ("axisname", C.c_char_p * 3), ("axislabel", C.c_char_p * 3),
                ("axisunits", C.c_char_p * 3), ("axismin", C.c_double * 3),
                ("axismax", C.c_double * 3)]


def io():
    """Loads libqpio.so; fails loudly when it has not been built."""
    global _io
    if _io is None:
        if not os.path.exists(IOPATH):
            raise RuntimeError(f"{IOPATH} is missing: build it (make -C quickpic-opensource_amd)")
        L = C.CDLL(IOPATH)
        L.qpio_last_error.restype = C.c_char_p
        P = C.POINTER(QpioFile)
        L.qpio_path.argtypes = [P, C.c_char_p, C.c_size_t]
        L.qpio_write_grid.argtypes = [P, C.c_int, C.POINTER(C.c_int64), C.c_void_p]
        L.qpio_write_raw.argtypes = [P, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
        L.qpio_write_rst.argtypes = [P, C.c_void_p, C.c_int64]
        L.qpio_read_rst.argtypes = [P, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        _io = L
    return _io


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"qpio: {io().qpio_last_error().decode()}")


@dataclass
class H5File:
    """type hdf5file (hdf5io_class.f03:18-35) with its defaults."""
    filename: str = "file.h5"
    dataname: str = "Data"
    timeunits: str = "a.u."
    ty: str = "grid"
    units: str = "a.u."
    label: str = "Data"
    n: int = 1
    rank: int = 2
    t: float = 1.0
    dt: float = 1.0
    axisname: tuple = ("x1", "x2", "x3")
    axislabel: tuple = ("x1", "x2", "x3")
    axisunits: tuple = ("a.u.", "a.u.", "a.u.")
    axismin: tuple = (0.0, 0.0, 0.0)
    axismax: tuple = (1.0, 1.0, 1.0)

    def struct(self):
        s = QpioFile()
        enc = lambda v: v.encode()   # noqa: E731
        s.filename, s.dataname = enc(self.filename), enc(self.dataname)
        s.timeunits, s.ty = enc(self.timeunits), enc(self.ty)
        s.units, s.label = enc(self.units), enc(self.label)
        s.n, s.rank, s.t, s.dt = self.n, self.rank, self.t, self.dt
        for i in range(3):
            s.axisname[i] = enc(self.axisname[i])
            s.axislabel[i] = enc(self.axislabel[i])
            s.axisunits[i] = enc(self.axisunits[i])
            s.axismin[i] = self.axismin[i]
            s.axismax[i] = self.axismax[i]
        return s

    def path(self):
        buf = C.create_string_buffer(4096)
        st = self.struct()
        _chk(io().qpio_path(C.byref(st), buf, 4096))
        return buf.value.decode()


def write_grid(f: H5File, data):
    """data: C-ordered (nz, ny, nx) or (n2, n1): written with Fortran
    dimensions reversed, as the reference's Fortran arrays."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    gs = (C.c_int64 * a.ndim)(*reversed(a.shape))
    st = f.struct()
    os.makedirs(os.path.dirname(f.path()) or ".", exist_ok=True)
    _chk(io().qpio_write_grid(C.byref(st), a.ndim, gs, a.ctypes.data))


def write_raw(f: H5File, part, dspl, delta, origin):
    a = np.ascontiguousarray(part, dtype=np.float64)
    d = np.ascontiguousarray(delta, dtype=np.float64)
    o = np.ascontiguousarray(origin, dtype=np.float64)
    st = f.struct()
    os.makedirs(os.path.dirname(f.path()) or ".", exist_ok=True)
    _chk(io().qpio_write_raw(C.byref(st), a.ctypes.data, a.shape[0], dspl, d.ctypes.data,
                             o.ctypes.data))


def write_rst(f: H5File, part):
    a = np.ascontiguousarray(part, dtype=np.float64)
    st = f.struct()
    os.makedirs(os.path.dirname(f.path()) or ".", exist_ok=True)
    _chk(io().qpio_write_rst(C.byref(st), a.ctypes.data, a.shape[0]))


def read_rst(f: H5File, max_records):
    a = np.zeros((max(1, max_records), 7))
    n = C.c_int64()
    st = f.struct()
    _chk(io().qpio_read_rst(C.byref(st), a.ctypes.data, max_records, C.byref(n)))
    return a[:n.value].copy()


# name -> (directory name, dataset name, units, label, component)
# (simulation_class.f03:1106-1168 for fields, :985-1020 for species)
FIELD_DIAGS = {
    "ex": ("Ex", "ex", "mc\\omega_p/e", "Electric Field", 1),
    "ey": ("Ey", "ey", "mc\\omega_p/e", "Electric Field", 2),
    "ez": ("Ez", "ez", "mc\\omega_p/e", "Electric Field", 3),
    "bx": ("Bx", "bx", "mc\\omega_p/e", "Magnetic Field", 1),
    "by": ("By", "by", "mc\\omega_p/e", "Magnetic Field", 2),
    "bz": ("Bz", "bz", "mc\\omega_p/e", "Magnetic Field", 3),
    "psi": ("Psi", "psi", "mc^2", "\\Psi", 1),
}
SPECIES_DIAGS = {
    "charge": ("Charge", "charge", "n_0", "Charge Density", 1),
    "jx": ("Jx", "jx", "n_0 c", "J_x", 1),
    "jy": ("Jy", "jy", "n_0 c", "J_y", 2),
    "jz": ("Jz", "jz", "n_0 c", "J_z", 3),
}
SLICE_CODE = {"yz": 1, "xz": 2, "xy": 3}


@dataclass
class Diag:
    """One entry of the reference's diag list: every df steps, component dim
    of source (a callable giving the Field3d) as a 3-D grid or a slice."""
    df: int
    file: H5File
    source: object = None
    dim: int = 1
    slice: int = 0
    slice_pos: int = 0
    kind: str = "grid"          # grid | raw | rst
    beam: int = -1
    sample: int = 1
    extra: dict = field(default_factory=dict)


class Diagnostics:
    """The diag list of a deck run (init_diag) and its dump (diag_simulation)."""

    def __init__(self, run, deck: dict, root: str = "."):
        self.run = run
        cfg = run.cfg
        self.root = root
        self._tstep = None
        xmin, xmax, ymin, ymax, zmin, zmax = cfg.box
        self.box = (xmin, xmax, ymin, ymax, zmin, zmax)
        self.dt = cfg.dt
        self.diags: list[Diag] = []
        self.psi3d = self.cu3d = None
        self.species_q3 = {}
        sim, ctx = run.sim, run.ctx
        for i in range(len(cfg.beams)):
            for d in deck["beam"][i].get("diag", []):
                nd = int(d.get("ndump", 0))
                if nd <= 0:
                    continue
                for name in d.get("name", []):
                    if name == "charge":
                        src = self._beam_charge_source(i)
                        self._grid_entries(nd, d, src, 1, f"Beam{i + 1:04d}", "Charge",
                                           "charge", "n_0", "Charge Density", "_slice_",
                                           "charge_slice_")
                    elif name == "raw":
                        f = H5File(filename=self._dir(f"Beam{i + 1:04d}/Raw/"), dataname="raw",
                                   units="", label="Beam Raw", ty="particles",
                                   timeunits="1 / \\omega_p", dt=self.dt)
                        self.diags.append(Diag(nd, f, kind="raw", beam=i,
                                               sample=int(d.get("sample", 1))))
        for i in range(len(cfg.species)):
            for d in deck["species"][i].get("diag", []):
                nd = int(d.get("ndump", 0))
                if nd <= 0:
                    continue
                for name in d.get("name", []):
                    if name not in SPECIES_DIAGS:
                        continue
                    sn1, sn2, sn3, sn4, dim = SPECIES_DIAGS[name]
                    if name == "charge":
                        if i not in self.species_q3:
                            self.species_q3[i] = qps.Field3d(ctx, 1)
                            sim.species_q3(i, self.species_q3[i])
                        src = (lambda k=i: self.species_q3[k])
                    else:
                        if self.cu3d is None:
                            self.cu3d = qps.Field3d(ctx, 3)
                            sim.diag3d(qps.DIAG_CU, self.cu3d)
                        src = (lambda: self.cu3d)
                    self._grid_entries(nd, d, src, dim, f"Species{i + 1:04d}", sn1, sn2, sn3, sn4,
                                       "_slice_", sn2 + "_slice_")
        for d in deck.get("field", {}).get("diag", []):
            nd = int(d.get("ndump", 0))
            if nd <= 0:
                continue
            for name in d.get("name", []):
                if name not in FIELD_DIAGS:
                    continue
                sn1, sn2, sn3, sn4, dim = FIELD_DIAGS[name]
                if name == "psi":
                    if self.psi3d is None:
                        self.psi3d = qps.Field3d(ctx, 1)
                        sim.diag3d(qps.DIAG_PSI, self.psi3d)
                    src = (lambda: self.psi3d)
                elif name.startswith("e"):
                    src = (lambda: run.sim.bexyz)
                else:
                    src = (lambda: run.sim.bbxyz)
                self._grid_entries(nd, d, src, dim, "Fields", sn1, sn2, sn3, sn4, "_slice",
                                   sn2 + "slice", field_dirs=True)
        s = deck["simulation"]
        if bool(s.get("dump_restart", False)):
            nd = int(s["ndump_restart"])
            for i in range(len(cfg.beams)):
                f = H5File(filename=self._dir(f"RST/Beam-{i + 1:04d}/"),
                           dataname=f"RST-beam{i + 1:04d}-{0:010d}", ty="rst")
                self.diags.append(Diag(nd, f, kind="rst", beam=i))

    def _dir(self, rel):
        return os.path.join(self.root, rel)

    def _beam_charge_source(self, i):
        """The beam's 3-D charge as a Field3d-like object with plane()."""
        run = self.run

        diags = self

        class _Q:
            dim = 1

            def __init__(self):
                self.vol, self.at = None, None

            def plane(self, comp, lpos):
                if self.at != diags._tstep:      # one download per dump
                    self.vol, self.at = run.beams[i].charge(), diags._tstep
                return self.vol[lpos - 1]
        q = _Q()
        return lambda: q

    def _grid_entries(self, nd, d, src, dim, top, sn1, sn2, sn3, sn4, sdir, sname,
                      field_dirs=False):
        xmin, xmax, ymin, ymax, zmin, zmax = self.box
        common = dict(timeunits="1 / \\omega_p", dt=self.dt,
                      axisunits=("c / \\omega_p",) * 3, units=sn3, label=sn4)
        if "slice" in d:
            for k, (plane, pos) in enumerate(d["slice"], start=1):
                if plane == "yz":
                    ax = dict(axismin=(ymin, zmin, xmin), axismax=(ymax, zmax, xmax),
                              axisname=("y", "\\xi", "x"), axislabel=("y", "\\xi", "x"))
                elif plane == "xz":
                    ax = dict(axismin=(xmin, zmin, ymin), axismax=(xmax, zmax, ymax),
                              axisname=("x", "\\xi", "y"), axislabel=("x", "\\xi", "y"))
                else:
                    ax = dict(axismin=(xmin, ymin, zmin), axismax=(xmax, ymax, zmax),
                              axisname=("x", "y", "z"), axislabel=("x", "y", "z"))
                rel = f"{top}/{sn1}{sdir}{k:04d}/"
                f = H5File(filename=self._dir(rel), dataname=sname + plane, rank=2,
                           **common, **ax)
                self.diags.append(Diag(nd, f, src, dim, SLICE_CODE[plane], int(pos)))
        else:
            f = H5File(filename=self._dir(f"{top}/{sn1}/"), dataname=sn2, rank=3,
                       axisname=("x", "y", "\\xi"), axislabel=("x", "y", "\\xi"),
                       axismin=(xmin, ymin, zmin), axismax=(xmax, ymax, zmax), **common)
            self.diags.append(Diag(nd, f, src, dim))

    # ---- data --------------------------------------------------------------
    def grid_data(self, f3, dim, sl, pos):
        """The reference's written array of component dim of a 3-D field
        (ufield3d_class.f03:492-600): nodes 0..n-1 of x and y, planes 1..nz;
        slices select one x node (yz), y node (xz) or plane (xy)."""
        ctx = self.run.ctx
        nx, ny, nz = ctx.nx, ctx.ny, ctx.nz
        if sl == SLICE_CODE["xy"]:
            return f3.plane(dim, pos)[:ny, :nx]
        out = np.zeros((nz, ny if sl == SLICE_CODE["yz"] else nx) if sl else (nz, ny, nx))
        for lpos in range(1, nz + 1):
            p = f3.plane(dim, lpos)
            if sl == SLICE_CODE["yz"]:
                out[lpos - 1] = p[:ny, pos - 1]
            elif sl == SLICE_CODE["xz"]:
                out[lpos - 1] = p[pos - 1, :nx]
            else:
                out[lpos - 1] = p[:ny, :nx]
        return out

    def dump(self, tstep):
        """diag_simulation: every entry whose period divides tstep."""
        written = []
        self._tstep = tstep
        origin = (self.box[0], self.box[2], self.box[4])
        delta = (self.run.ctx.dex, self.run.ctx.dex, self.run.ctx.dxi)
        for d in self.diags:
            if tstep % d.df != 0:
                continue
            d.file.n, d.file.t = tstep, tstep * self.dt
            if d.kind == "raw":
                write_raw(d.file, self.run.beams[d.beam].download(), d.sample, delta, origin)
            elif d.kind == "rst":
                write_rst(d.file, self.run.beams[d.beam].download())
            else:
                write_grid(d.file, self.grid_data(d.source(), d.dim, d.slice, d.slice_pos))
            written.append(d.file.path())
        return written


def read_restart(run, deck, root="."):
    """rpart into every beam (simulation_class.f03:459-471) for a deck with
    read_restart: the step in restart_timestep."""
    s = deck["simulation"]
    step = int(s["restart_timestep"])
    for i, b in enumerate(run.beams):
        f = H5File(filename=os.path.join(root, f"RST/Beam-{i + 1:04d}/"),
                   dataname=f"RST-beam{i + 1:04d}-{0:010d}", n=step, ty="rst")
        cap = max(b.np, int(getattr(run.cfg.beams[i], "npmax", 0) or 0))
        part = read_rst(f, max(cap, 1) * 4)
        b.upload(part)
    return step
