"""Python host mirror of libqpslice.so (the MI355X quasi-static slice solver).

Thin ctypes classes over the C ABI of include/qpslice.h, named after the
reference's Fortran classes so that a test or driver reads like the
reference's host code:

    Ctx       -- spect2d/spect3d + the box (simulation_class.f03:151-224)
    Field2d   -- field2d (field2d_class.f03): as, fftrk, fftkr, pot, grad, div,
                 bf, bfqp, elf, smooth, add, sub, mult, ag, cg, cb, ca
    Field3d   -- field3d (ufield3d_class.f03:396-490)
    Species   -- species2d (species2d_class.f03): qdp, extpsi, amjdp, push,
                 cbq, renew
    Beam      -- beam3d (beam3d_class.f03): qdp, qdpcopy, push
    Sim       -- the 3-D step and xi-slice loop (simulation_class.f03:611-765)

Every call raises QPError when the library returns a nonzero status, the
counterpart of the reference's perrors%equit.  numpy arrays exchanged with the
library use these shapes (float64):
    real field2d (dim, ny+1, nx+1), k field2d (dim, nx+1, ny+1),
    field3d (dim, nz+1, ny+1, nx+1), plasma particles (np, 8), beam (np, 7).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libqpslice.so")

FP64, FP32 = 0, 1
(PROF_FFT, PROF_TRANSPOSE, PROF_KSPACE, PROF_ELEMENTWISE, PROF_DEPOSIT_Q, PROF_EXTPSI,
 PROF_AMJDP, PROF_PUSH, PROF_SORT, PROF_BORDER, PROF_BEAM_DEPOSIT, PROF_BEAM_PUSH) = range(12)
PROF_NAMES = ("fft", "transpose", "kspace", "elementwise", "deposit_q", "extpsi", "amjdp",
              "push", "sort", "border", "beam_deposit", "beam_push")
(TUNE_DEPOSIT_Q, TUNE_AMJDP, TUNE_SPECULATE, TUNE_SEED, TUNE_FFT, TUNE_PUSH, TUNE_MIX,
 TUNE_STRICT) = (0, 1, 2, 3, 4, 5, 6, 7)
(F_QB, F_QE, F_PSIT, F_PSI, F_DIVVPOT, F_FXY, F_BXYZ, F_CU, F_DCU, F_AMU,
 F_EPW, F_EPWB) = range(12)
DIAG_PSI, DIAG_CU = 0, 1


class QPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qpslice error {code}: {msg}")
        self.code = code


_lib = None
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_I64 = C.c_int64


class SpeciesParams(C.Structure):
    _fields_ = [("ppc", _I * 2), ("density", _D), ("q", _D), ("m", _D)]


class SpeciesProfile(C.Structure):
    _fields_ = [("profile", _I), ("center", _D * 2), ("inner_radius", _D * 2),
                ("outer_radius", _D * 2), ("radius", _D), ("width", _D), ("nr", _I),
                ("r", C.POINTER(_D)), ("fr", C.POINTER(_D))]


class BeamParams(C.Structure):
    _fields_ = [("center", _D * 3), ("sigma", _D * 3), ("sigma_v", _D * 3),
                ("centroid_x", _D * 3), ("centroid_y", _D * 3),
                ("gamma", _D), ("peak", _D), ("q", _D), ("m", _D),
                ("np", _I * 3), ("quiet", _I), ("evolution", _I), ("npmax", _I64)]


class BeamFileParams(C.Structure):
    _fields_ = [("file_name", C.c_char_p), ("center", _D * 3), ("n0", _D), ("q", _D),
                ("m", _D), ("np", _I), ("evolution", _I), ("npmax", _I64)]


class BeamTorusParams(C.Structure):
    _fields_ = [("center", _D * 3), ("range_x", _D * 2), ("range_y", _D * 2),
                ("range_z", _D * 2), ("ppc", _I * 3), ("q", _D), ("m", _D), ("np", _D),
                ("gamma", _D), ("r1", _D), ("r2", _D), ("evolution", _I), ("npmax", _I64)]


class BeamTwissParams(C.Structure):
    _fields_ = [("center", _D * 3), ("alpha", _D * 2), ("beta", _D * 2), ("emittance", _D * 2),
                ("sigma_z", _D), ("sigma_vz", _D), ("gamma", _D), ("peak", _D), ("q", _D),
                ("m", _D), ("np", _I * 3), ("quiet", _I), ("evolution", _I), ("npmax", _I64)]


def lib():
    """Loads libqpslice.so; fails loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it (make -C quickpic-opensource_amd)")
        L = C.CDLL(LIBPATH)
        L.qp_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _call(fn, *args):
    rc = fn(*args)
    if rc != 0:
        raise QPError(rc, lib().qp_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(_P)


def _ints(v):
    return (C.c_int * len(v))(*v)


def _dbls(v):
    return (C.c_double * len(v))(*v)


class Ctx:
    def __init__(self, indx, indy, indz, box=(-10.0, 10.0, -10.0, 10.0, 0.0, 10.0),
                 precision=FP64, device=0):
        self.h = _P()
        _call(lib().qp_ctx_create, indx, indy, indz, _dbls(box), precision, device,
              C.byref(self.h))
        nx, ny, nz, pr = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        dex, dxi = C.c_double(), C.c_double()
        _call(lib().qp_ctx_info, self.h, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(pr),
              C.byref(dex), C.byref(dxi))
        self.nx, self.ny, self.nz = nx.value, ny.value, nz.value
        self.precision = pr.value
        self.dex, self.dxi = dex.value, dxi.value
        self._children = []

    def _adopt(self, obj):
        import weakref
        self._children.append(weakref.ref(obj))

    def sync(self):
        _call(lib().qp_ctx_sync, self.h)

    def stream(self):
        """The context's HIP stream handle (an int), on which the library
        enqueues every call."""
        h = C.c_void_p()
        _call(lib().qp_ctx_stream, self.h, C.byref(h))
        return h.value

    def profile(self, enable=True, classes=None):
        """Starts (and resets) or stops kernel-class timing with HIP events;
        classes: the PROF_* kinds to time (default: all)."""
        if classes is None:
            _call(lib().qp_ctx_profile, self.h, 1 if enable else 0)
        else:
            mask = sum(1 << k for k in classes) if enable else 0
            _call(lib().qp_ctx_profile_classes, self.h, C.c_uint(mask))

    def profile_every(self, every):
        """Times one launch in `every` of each enabled class (1: all)."""
        _call(lib().qp_ctx_profile_every, self.h, int(every))

    def profile_get(self, kind):
        """(total milliseconds, launches, algorithmic bytes) of one kernel
        class (PROF_*)."""
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        _call(lib().qp_ctx_profile_get, self.h, kind, C.byref(ms), C.byref(n), C.byref(b))
        return ms.value, n.value, b.value

    def tune(self, knob, value):
        """Selects variant `value` of a tuning knob (TUNE_*); -1 is the
        default.  Variants change launch shapes, kernel forms or the order of
        per-cell sums, never what is computed."""
        _call(lib().qp_ctx_tune, self.h, knob, value)

    def tune_get(self, knob):
        """(current value, number of values) of a tuning knob."""
        v, n = C.c_int(), C.c_int()
        _call(lib().qp_ctx_tune_get, self.h, knob, C.byref(v), C.byref(n))
        return v.value, n.value

    def close(self):
        """Destroys the context after every object created on it."""
        if self.h:
            for r in reversed(self._children):
                o = r()
                if o is not None:
                    o.close()
            self._children = []
            lib().qp_ctx_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field2d:
    def __init__(self, ctx: Ctx, dim, state=0, handle=None):
        self.ctx, self.dim = ctx, dim
        self.owned = handle is None
        if handle is None:
            self.h = _P()
            _call(lib().qp_field2d_create, ctx.h, dim, state, C.byref(self.h))
            ctx._adopt(self)
        else:
            self.h = handle

    @property
    def state(self):
        s, d = C.c_int(), C.c_int()
        _call(lib().qp_field2d_state, self.h, C.byref(s), C.byref(d))
        return s.value

    def upload(self, arr, space=None):
        """arr: real (dim, ny+1, nx+1) or k (dim, nx+1, ny+1)."""
        c = self.ctx
        if space is None:
            space = 0 if arr.shape[1:] == (c.ny + 1, c.nx + 1) else 1
        a = np.ascontiguousarray(np.transpose(np.asarray(arr, np.float64), (1, 2, 0)))
        ld1 = a.shape[1]
        ld2 = a.shape[0]
        _call(lib().qp_field2d_upload, self.h, space, _ptr(a), ld1, ld2)

    def device(self, space):
        """(device address, row pitch, component stride) of one space, in
        elements of the context's precision."""
        p, ld, cs = C.c_void_p(), C.c_int(), C.c_size_t()
        _call(lib().qp_field2d_device, self.h, space, C.byref(p), C.byref(ld), C.byref(cs))
        return p.value, ld.value, cs.value

    def download(self, space=None):
        c = self.ctx
        space = self.state if space is None else space
        if space == 0:
            a = np.zeros((c.ny + 1, c.nx + 1, self.dim))
        else:
            a = np.zeros((c.nx + 1, c.ny + 1, self.dim))
        _call(lib().qp_field2d_download, self.h, space, _ptr(a), a.shape[1], a.shape[0])
        return np.ascontiguousarray(np.transpose(a, (2, 0, 1)))

    def as_(self, v):
        _call(lib().qp_field2d_as, self.h, _D(v))

    def asa(self, that):
        _call(lib().qp_field2d_asa, self.h, that.h)

    def fftrk(self, kind):
        _call(lib().qp_field2d_fftrk, self.h, kind)

    def fftkr(self, kind):
        _call(lib().qp_field2d_fftkr, self.h, kind)

    def div(self, out):
        _call(lib().qp_field2d_div, self.h, out.h)

    def grad(self, out):
        _call(lib().qp_field2d_grad, self.h, out.h)

    def pot(self, out):
        _call(lib().qp_field2d_pot, self.h, out.h)

    def smooth(self, out):
        _call(lib().qp_field2d_smooth, self.h, out.h)

    def elf(self, out):
        _call(lib().qp_field2d_elf, self.h, out.h)

    def bf(self, out):
        _call(lib().qp_field2d_bf, self.h, out.h)

    def bfqp(self, dcu, amu, bxy, c, dex):
        _call(lib().qp_field2d_bfqp, self.h, dcu.h, amu.h, bxy.h, _D(c), _D(dex))

    def add(self, a1, a2, dim=None, dim1=None, dim2=None):
        if dim is None:
            _call(lib().qp_field2d_add, self.h, a1.h, a2.h, 0, None, None, None)
        else:
            _call(lib().qp_field2d_add, self.h, a1.h, a2.h, len(dim), _ints(dim), _ints(dim1),
                  _ints(dim2))

    def sub(self, a1, a2, dim=None, dim1=None, dim2=None):
        if dim is None:
            _call(lib().qp_field2d_sub, self.h, a1.h, a2.h, 0, None, None, None)
        else:
            _call(lib().qp_field2d_sub, self.h, a1.h, a2.h, len(dim), _ints(dim), _ints(dim1),
                  _ints(dim2))

    def mult(self, a1, value, dim=None, dim1=None):
        if dim is None:
            _call(lib().qp_field2d_mult, self.h, a1.h, 0, None, None, _dbls([value]))
        else:
            _call(lib().qp_field2d_mult, self.h, a1.h, len(dim), _ints(dim), _ints(dim1),
                  _dbls(value))

    def ag(self):
        _call(lib().qp_field2d_ag, self.h)

    def cg(self):
        _call(lib().qp_field2d_cg, self.h)

    def cb(self, f3, lpos, sdim, ddim):
        _call(lib().qp_field2d_cb, self.h, f3.h, lpos, len(sdim), _ints(sdim), _ints(ddim))

    def ca(self, f3, lpos, sdim, ddim):
        _call(lib().qp_field2d_ca, self.h, f3.h, lpos, len(sdim), _ints(sdim), _ints(ddim))

    def close(self):
        if self.owned and self.h:
            lib().qp_field2d_destroy(self.h)
        self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field3d:
    def __init__(self, ctx: Ctx, dim, handle=None):
        self.ctx, self.dim = ctx, dim
        self.owned = handle is None
        if handle is None:
            self.h = _P()
            _call(lib().qp_field3d_create, ctx.h, dim, C.byref(self.h))
            ctx._adopt(self)
        else:
            self.h = handle

    def as_(self, v):
        _call(lib().qp_field3d_as, self.h, _D(v))

    def upload(self, arr):
        a = np.ascontiguousarray(np.transpose(np.asarray(arr, np.float64), (1, 2, 3, 0)))
        _call(lib().qp_field3d_upload, self.h, _ptr(a), a.shape[2], a.shape[1], a.shape[0])

    def download(self):
        c = self.ctx
        a = np.zeros((c.nz + 1, c.ny + 1, c.nx + 1, self.dim))
        _call(lib().qp_field3d_download, self.h, _ptr(a), a.shape[2], a.shape[1], a.shape[0])
        return np.ascontiguousarray(np.transpose(a, (3, 0, 1, 2)))

    def plane(self, comp, lpos):
        c = self.ctx
        a = np.zeros((c.ny + 1, c.nx + 1))
        _call(lib().qp_field3d_download_plane, self.h, comp, lpos, _ptr(a), c.nx + 1, c.ny + 1)
        return a

    def close(self):
        if self.owned and self.h:
            lib().qp_field3d_destroy(self.h)
        self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Species:
    def __init__(self, ctx: Ctx, ppc=(2, 2), density=1.0, q=-1.0, m=1.0, dt=None, s=0.0):
        self.ctx = ctx
        p = SpeciesParams()
        p.ppc[0], p.ppc[1] = ppc
        p.density, p.q, p.m = density, q, m
        self.h = _P()
        _call(lib().qp_species_create, ctx.h, C.byref(p), _D(ctx.dxi if dt is None else dt),
              _D(s), C.byref(self.h))
        ctx._adopt(self)

    def renew(self, s=0.0):
        _call(lib().qp_species_renew, self.h, _D(s))

    def set_profile(self, profile, center=(0.0, 0.0), inner_radius=(0.0, 0.0),
                    outer_radius=(0.0, 0.0), radius=0.0, width=1.0, radial=None, s_now=0.0):
        """Transverse profile 10 (elliptical annulus), 11 (Gaussian ring) or
        12 (radial = (r points, densities)), fdist2d_class.f03:318-950;
        reloads the plasma at s_now.  0: uniform."""
        p = SpeciesProfile()
        p.profile = int(profile)
        p.center[0], p.center[1] = center
        p.inner_radius[0], p.inner_radius[1] = inner_radius
        p.outer_radius[0], p.outer_radius[1] = outer_radius
        p.radius, p.width = radius, width
        keep = None
        if radial is not None:
            r = np.ascontiguousarray(radial[0], dtype=np.float64)
            fr = np.ascontiguousarray(radial[1], dtype=np.float64)
            if r.shape != fr.shape:
                raise QPError(-1, "dist2d_012: The piecewise_radial_density and r array have "
                                  "different sizes!")
            keep = (r, fr)
            p.nr = len(r)
            p.r = r.ctypes.data_as(C.POINTER(_D))
            p.fr = fr.ctypes.data_as(C.POINTER(_D))
        _call(lib().qp_species_set_profile, self.h, C.byref(p), _D(s_now))
        del keep

    def set_longitudinal(self, s_pts, fs, s_now=0.0):
        """Piecewise longitudinal density (fdist2d_class.f03:190-201,
        233-251); reloads the plasma at s_now.  Empty lists: uniform."""
        n = len(s_pts)
        _call(lib().qp_species_set_longitudinal, self.h, n, _dbls(list(s_pts) or [0.0]),
              _dbls(list(fs) or [0.0]), _D(s_now))

    def qdp(self, q):
        _call(lib().qp_species_qdp, self.h, q.h)

    def qdp_noalloc(self, q):
        """qdp with the deposit's scratch in the library's static array (no
        allocation; grids up to 4096 tiles in FP64, 8192 in FP32)."""
        _call(lib().qp_species_qdp_noalloc, self.h, q.h)

    def extpsi(self, psi, dex):
        _call(lib().qp_species_extpsi, self.h, psi.h, _D(dex))

    def amjdp(self, ef, bf, psit, cu, amu, dcu, dex):
        _call(lib().qp_species_amjdp, self.h, ef.h, bf.h, psit.h, cu.h, amu.h, dcu.h, _D(dex))

    def push(self, ef, bf, psit, dex):
        _call(lib().qp_species_push, self.h, ef.h, bf.h, psit.h, _D(dex))

    def cbq(self, q3, pos):
        _call(lib().qp_species_cbq, self.h, q3.h if q3 is not None else None, pos)

    @property
    def np(self):
        n = C.c_int64()
        _call(lib().qp_species_count, self.h, C.byref(n))
        return n.value

    def download(self):
        n = self.np
        a = np.zeros((n, 8))
        _call(lib().qp_species_download, self.h, _ptr(a), _I64(n))
        return a

    def upload(self, part):
        a = np.ascontiguousarray(part, np.float64)
        _call(lib().qp_species_upload, self.h, _ptr(a), _I64(a.shape[0]), _I64(a.shape[0]))

    def cell_counts(self):
        c = self.ctx
        a = np.zeros(c.nx * c.ny, np.int32)
        _call(lib().qp_species_cell_counts, self.h, _ptr(a))
        return a

    def tile_counts(self):
        c = self.ctx
        a = np.zeros((c.nx // 16) * (c.ny // 16), np.int32)
        _call(lib().qp_species_tile_counts, self.h, _ptr(a))
        return a

    def close(self):
        if self.h:
            lib().qp_species_destroy(self.h)
        self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Beam:
    def __init__(self, ctx: Ctx, center=(0.0, 0.0, 2.0), sigma=(0.5, 0.5, 0.5),
                 sigma_v=(1.0, 1.0, 0.0), np_=(64, 64, 256), gamma=2.0e4, peak=4.0,
                 q=-1.0, m=1.0, quiet=True, evolution=True, dt=10.0, seed=0, npmax=0,
                 centroid_x=(0.0, 0.0, 0.0), centroid_y=(0.0, 0.0, 0.0)):
        self.ctx = ctx
        p = BeamParams()
        for i in range(3):
            p.center[i], p.sigma[i], p.sigma_v[i] = center[i], sigma[i], sigma_v[i]
            p.centroid_x[i], p.centroid_y[i] = centroid_x[i], centroid_y[i]
            p.np[i] = np_[i]
        p.gamma, p.peak, p.q, p.m = gamma, peak, q, m
        p.quiet = 1 if quiet else 0
        p.evolution = 1 if evolution else 0
        p.npmax = npmax
        self.h = _P()
        _call(lib().qp_beam_create, ctx.h, C.byref(p), _D(dt), C.c_uint64(seed), C.byref(self.h))
        ctx._adopt(self)

    def qdp(self):
        _call(lib().qp_beam_qdp, self.h)

    def qdpcopy(self, q, slice_):
        _call(lib().qp_beam_qdpcopy, self.h, q.h, slice_)

    def push(self, ef, bf, dex, dez):
        _call(lib().qp_beam_push, self.h, ef.h, bf.h, _D(dex), _D(dez))

    @property
    def np(self):
        n = C.c_int64()
        _call(lib().qp_beam_count, self.h, C.byref(n))
        return n.value

    def download(self):
        n = self.np
        a = np.zeros((n, 7))
        _call(lib().qp_beam_download, self.h, _ptr(a), _I64(n))
        return a

    def upload(self, part):
        a = np.ascontiguousarray(part, np.float64)
        _call(lib().qp_beam_upload, self.h, _ptr(a), _I64(a.shape[0]), _I64(a.shape[0]))

    def charge(self):
        c = self.ctx
        a = np.zeros((c.nz + 1, c.ny + 1, c.nx + 1))
        _call(lib().qp_beam_charge_download, self.h, _ptr(a), c.nx + 1, c.ny + 1, c.nz + 1)
        return a

    def moments(self):
        g, q = C.c_double(), C.c_double()
        _call(lib().qp_beam_moments, self.h, C.byref(g), C.byref(q))
        return g.value, q.value

    def close(self):
        if self.h:
            lib().qp_beam_destroy(self.h)
        self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TwissBeam(Beam):
    """Beam profile 2 (fdist3d_002): Twiss transverse phase space, Gaussian
    in z.  Lengths in c/wp as in the input deck."""

    def __init__(self, ctx: Ctx, center=(0.0, 0.0, 2.0), alpha=(0.0, 0.0), beta=(1.0, 1.0),
                 emittance=(1e-4, 1e-4), sigma_z=0.5, sigma_vz=0.0, np_=(64, 64, 256),
                 gamma=2.0e4, peak=4.0, q=-1.0, m=1.0, quiet=True, evolution=True, dt=10.0,
                 seed=0, npmax=0):
        self.ctx = ctx
        p = BeamTwissParams()
        for i in range(3):
            p.center[i], p.np[i] = center[i], np_[i]
        for i in range(2):
            p.alpha[i], p.beta[i], p.emittance[i] = alpha[i], beta[i], emittance[i]
        p.sigma_z, p.sigma_vz = sigma_z, sigma_vz
        p.gamma, p.peak, p.q, p.m = gamma, peak, q, m
        p.quiet = 1 if quiet else 0
        p.evolution = 1 if evolution else 0
        p.npmax = npmax
        self.h = _P()
        _call(lib().qp_beam_create_twiss, ctx.h, C.byref(p), _D(dt), C.c_uint64(seed),
              C.byref(self.h))
        ctx._adopt(self)


class FileBeam(Beam):
    """Beam profile 3 (fdist3d_003): np particles read from a text file of
    records "z x y vz vx vy q" in units of 5.32150254e9/sqrt(n0)."""

    def __init__(self, ctx: Ctx, file_name, np_, center=(0.0, 0.0, 0.0), n0=1e16, q=-1.0,
                 m=1.0, evolution=True, dt=10.0, npmax=0):
        self.ctx = ctx
        p = BeamFileParams()
        p.file_name = os.fsencode(file_name)
        for i in range(3):
            p.center[i] = center[i]
        p.n0, p.q, p.m, p.np = n0, q, m, int(np_)
        p.evolution = 1 if evolution else 0
        p.npmax = npmax
        self.h = _P()
        _call(lib().qp_beam_create_file, ctx.h, C.byref(p), _D(dt), C.byref(self.h))
        ctx._adopt(self)


class TorusBeam(Beam):
    """Beam profile 100 (fdist3d_100): a quiet-start Gaussian torus (radius
    r1, width r2) over the range box.  Lengths in c/wp."""

    def __init__(self, ctx: Ctx, center=(0.0, 0.0, 2.0), range_x=(-3.0, 3.0),
                 range_y=(-3.0, 3.0), range_z=(1.0, 3.0), ppc=(2, 2, 2), q=-1.0, m=1.0,
                 np_=1.0, gamma=2.0e4, r1=1.0, r2=0.3, evolution=True, dt=10.0, npmax=0):
        self.ctx = ctx
        p = BeamTorusParams()
        for i in range(3):
            p.center[i], p.ppc[i] = center[i], ppc[i]
        for i in range(2):
            p.range_x[i], p.range_y[i], p.range_z[i] = range_x[i], range_y[i], range_z[i]
        p.q, p.m, p.np, p.gamma, p.r1, p.r2 = q, m, np_, gamma, r1, r2
        p.evolution = 1 if evolution else 0
        p.npmax = npmax
        self.h = _P()
        _call(lib().qp_beam_create_torus, ctx.h, C.byref(p), _D(dt), C.byref(self.h))
        ctx._adopt(self)


class Sim:
    def __init__(self, ctx: Ctx, species, beams, iter_=2):
        self.ctx = ctx
        self.species, self.beams = list(species), list(beams)
        sp = (_P * max(1, len(species)))(*[s.h for s in species])
        bm = (_P * max(1, len(beams)))(*[b.h for b in beams])
        self.h = _P()
        _call(lib().qp_sim_create, ctx.h, sp, len(species), bm, len(beams), iter_,
              C.byref(self.h))
        ctx._adopt(self)
        self.iter = iter_
        self._f = {}
        e, b = _P(), _P()
        _call(lib().qp_sim_field3d, self.h, C.byref(e), C.byref(b))
        self.bexyz = Field3d(ctx, 3, handle=e)
        self.bbxyz = Field3d(ctx, 3, handle=b)

    def field(self, which):
        if which not in self._f:
            dims = {F_QB: 1, F_QE: 1, F_PSIT: 1, F_PSI: 1, F_DIVVPOT: 1, F_FXY: 2, F_BXYZ: 3,
                    F_CU: 3, F_DCU: 2, F_AMU: 3, F_EPW: 2, F_EPWB: 2}
            h = _P()
            _call(lib().qp_sim_field2d, self.h, which, C.byref(h))
            self._f[which] = Field2d(self.ctx, dims[which], handle=h)
        return self._f[which]

    def species_q3(self, i, q3):
        _call(lib().qp_sim_species_q3, self.h, i, q3.h if q3 is not None else None)

    def diag3d(self, which, f3):
        """Keeps psi (DIAG_PSI, 1 component) or the current (DIAG_CU, 3) of
        every slice in the 3-D field f3 (None stops it)."""
        _call(lib().qp_sim_diag3d, self.h, which, f3.h if f3 is not None else None)

    def begin_step(self):
        _call(lib().qp_sim_begin_step, self.h)

    def slice(self, j, iter_max=None, tol=0.0):
        it, res = C.c_int(), C.c_double()
        _call(lib().qp_slice_run, self.h, j, self.iter if iter_max is None else iter_max,
              _D(tol), C.byref(it), C.byref(res))
        return it.value, res.value

    def prepare(self, j):
        """First part of slice j, up to the corrector seed."""
        _call(lib().qp_slice_prepare, self.h, j)

    def finish(self, j, iter_max=None, tol=0.0):
        """The rest of slice j after prepare; returns (passes, residual)."""
        it, res = C.c_int(), C.c_double()
        _call(lib().qp_slice_finish, self.h, j, self.iter if iter_max is None else iter_max,
              _D(tol), C.byref(it), C.byref(res))
        return it.value, res.value

    def residuals(self):
        """Corrector residuals of the last slice run with a tolerance, pass 1
        first."""
        a = np.zeros(64)
        n = C.c_int()
        _call(lib().qp_slice_residuals, self.h, _ptr(a), 64, C.byref(n))
        return a[:n.value].copy()

    def end_step(self, s=0.0):
        _call(lib().qp_sim_end_step, self.h, _D(s))

    def push_beams(self):
        """The first part of end_step: the beam push."""
        _call(lib().qp_sim_push_beams, self.h)

    def renew(self, s=0.0):
        """The second part of end_step: the plasma renewal at s."""
        _call(lib().qp_sim_renew, self.h, _D(s))

    def step(self, nslices=None):
        self.begin_step()
        n = self.ctx.nz if nslices is None else nslices
        for j in range(1, n + 1):
            self.slice(j)
        self.end_step()

    def close(self):
        for f in self._f.values():
            f.h = _P()
        if self.h:
            lib().qp_sim_destroy(self.h)
        self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scan:
    """One HIP graph holding one xi slice of each simulation (qp_scan_*):
    run(n) replays it n times, every simulation continuing from its own
    last slice.  Create it after each simulation's begin_step."""

    def __init__(self, sims, iter_max=2, tol=0.0):
        self.sims = list(sims)
        hs = (_P * len(self.sims))(*[x.h for x in self.sims])
        self.h = _P()
        _call(lib().qp_scan_create, hs, len(self.sims), iter_max, _D(tol), C.byref(self.h))
        for x in self.sims:   # closing any of the contexts closes the scan first
            x.ctx._adopt(self)

    def run(self, nslices):
        _call(lib().qp_scan_run, self.h, nslices)

    def passes(self, i, nslices):
        """Corrector passes of slices 1..nslices of simulation i."""
        a = np.zeros(nslices, dtype=np.int32)
        _call(lib().qp_scan_passes, self.h, i, a.ctypes.data_as(C.c_void_p), nslices)
        return a

    def close(self):
        if self.h:
            lib().qp_scan_destroy(self.h)
        self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exported_symbols(header=os.path.join(os.path.dirname(_HERE), "include", "qpslice.h")):
    """Names of the functions include/qpslice.h declares."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"\b(qp_[a-z0-9_]+)\s*\(", text)))
