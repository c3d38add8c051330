"""Pins the C oracle against the reference's own F77 kernels, compiled from
/root/reference by oracle/Makefile.ref (skipped where that build is absent,
e.g. on a machine without the reference tree).

Bars: the per-particle kernels and the k-space operators restate the
reference's operation order, so with identical inputs they agree to a few ulp
(the only difference is the summation order of the deposits: the reference
sums per 16x16 tile first).  The sine/cosine transforms are an independent
restatement (complex FFT of the symmetric extension instead of the
Numerical-Recipes recursion): they agree to 1e-13 relative.  ranorm and the
beam loader are bit-exact.
"""
import ctypes
import os
import shutil
import tempfile

import numpy as np
import pytest

from oracle import qpo
from oracle import refharness as rh

NX = NY = 64
DEX = 20.0 / NX
DXI = 10.0 / 128


def rel(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s


def fresh_ref():
    """A private copy of the reference library (its ranorm keeps SAVE state)."""
    d = tempfile.mkdtemp()
    p = os.path.join(d, "libqpref_copy.so")
    shutil.copy(rh.REF_SO, p)
    return rh.Ref(p)


@pytest.fixture(scope="module")
def plasma():
    part = qpo.plasma_init(NX, NY, 2, 2)
    qpo.order_by_tile(NX, NY, part)
    rng = np.random.default_rng(7)
    # give particles momenta and in-cell jitter so every branch is exercised
    part[:, 2:5] = rng.normal(0.0, 0.3, (part.shape[0], 3))
    return part


def smooth_fields(rng, dims, amp=0.05):
    """Band-limited random fields (sum of a few low sine/cosine modes)."""
    y, x = np.mgrid[0:NY + 1, 0:NX + 1]
    out = np.zeros((dims, NY + 1, NX + 1))
    for c in range(dims):
        for _ in range(4):
            kx, ky = rng.integers(1, 5, 2)
            out[c] += amp * rng.normal() * np.sin(np.pi * kx * x / NX + rng.uniform()) \
                * np.cos(np.pi * ky * y / NY + rng.uniform())
    return out


@pytest.mark.parametrize("kind,dim", sorted(rh.Ref.WRAP))
def test_transforms(ref, kind, dim):
    rng = np.random.default_rng(kind * 10 + dim)
    f = rng.standard_normal((dim, NY + 1, NX + 1))
    g = ref.kfield(dim, NX, NY)
    fr = ref.r_from_oracle(f, NX, NY)
    ref.transform(kind, fr, g, -1, NX, NY)
    assert rel(qpo.r2k(kind, f), ref.k_to_oracle(g)) < 1e-13
    k = rng.standard_normal((dim, NX + 1, NY + 1))
    g = ref.k_from_oracle(k)
    fr = ref.rfield(dim, NX, NY)
    ref.transform(kind, fr, g, 1, NX, NY)
    ref.copyguard(fr, NY)
    assert rel(qpo.k2r(kind, k), ref.r_to_oracle(fr, NX, NY)) < 1e-13


def test_pois_table(ref):
    g, s = qpo.pois_table(NX, NY)
    ffc = ref.pois_table(NX, NY)
    assert np.array_equal(g, ffc.real.T)
    assert np.array_equal(s, ffc.imag.T)


def _ks(rng, dim):
    return rng.standard_normal((dim, NX + 1, NY + 1))


def test_kspace_ops(ref):
    rng = np.random.default_rng(3)
    tab = qpo.pois_table(NX, NY)
    ffc = ref.pois_table(NX, NY)
    q = _ks(rng, 1)
    for name, fn, dout in (("pot", qpo.pot, 1), ("smooth", qpo.smooth, 1),
                           ("elf", qpo.elf, 2)):
        out = ref.kfield(dout, NX, NY)
        qin = ref.k_from_oracle(q)
        if dout == 1:
            o1 = np.asfortranarray(out[0])
            ref.kcall(name, ffc, NX, NY, np.asfortranarray(qin[0]), o1)
            out[0] = o1
        else:
            ref.kcall(name, ffc, NX, NY, np.asfortranarray(qin[0]), out)
        assert rel(fn(tab, q), ref.k_to_oracle(out)) < 1e-15, name
    cu = _ks(rng, 3)
    out = ref.kfield(3, NX, NY)
    ref.kcall("bf", ffc, NX, NY, ref.k_from_oracle(cu), out)
    assert rel(qpo.bf(tab, cu), ref.k_to_oracle(out)) < 1e-15
    f2 = _ks(rng, 2)
    out = ref.kfield(1, NX, NY)
    o1 = np.asfortranarray(out[0])
    ref.kcall("div", ffc, NX, NY, ref.k_from_oracle(f2), o1)
    assert rel(qpo.div(f2)[0], ref.k_to_oracle(o1[None])[0]) < 1e-15
    out = ref.kfield(2, NX, NY)
    ref.kcall("grad", ffc, NX, NY, np.asfortranarray(ref.k_from_oracle(q)[0]), out)
    assert rel(qpo.grad(q), ref.k_to_oracle(out)) < 1e-15


def test_bfqp(ref):
    rng = np.random.default_rng(4)
    tab = qpo.pois_table(NX, NY)
    ffc = ref.pois_table(NX, NY)
    cu, dcu, amu, b = _ks(rng, 3), _ks(rng, 2), _ks(rng, 3), _ks(rng, 2)
    rcu, rdcu, ramu, rb = (ref.k_from_oracle(a) for a in (cu, dcu, amu, b))
    qpo.bfqp(tab, cu, dcu, amu, b, DEX * DEX, DEX)
    ref.bfqp(ffc, rcu, rdcu, ramu, rb, DEX * DEX, DEX, NX, NY)
    assert rel(b, ref.k_to_oracle(rb)) < 1e-15
    assert rel(dcu, ref.k_to_oracle(rdcu)) < 1e-15


def test_charge_deposit(ref, plasma):
    part = plasma.copy()
    q = np.zeros((1, NY + 1, NX + 1))
    qpo.qdeposit(NX, NY, part, q)
    qpo.guard_fold(q)
    ppart, kpic, _ = ref.to_tiles(part, NX, NY)
    qr = ref.rfield(1, NX, NY)
    ref.qdeposit(ppart, kpic, qr, NX, NY)
    ref.fold(qr, NY)
    assert rel(q, ref.r_to_oracle(qr, NX, NY)) < 1e-14


def test_extpsi(ref, plasma):
    rng = np.random.default_rng(5)
    psi = smooth_fields(rng, 1, amp=0.5)
    part = plasma.copy()
    qpo.extpsi(part, psi, -1.0, DEX)
    ppart, kpic, _ = ref.to_tiles(plasma.copy(), NX, NY)
    ref.extpsi(ppart, kpic, ref.r_from_oracle(psi, NX, NY), -1.0, DEX, NX, NY)
    assert np.array_equal(part, ref.from_tiles(ppart, kpic))


def _fields(seed):
    rng = np.random.default_rng(seed)
    return smooth_fields(rng, 2), smooth_fields(rng, 3), smooth_fields(rng, 1)


def test_amjdp(ref, plasma):
    fxy, bxy, psit = _fields(6)
    part = plasma.copy()
    qpo.extpsi(part, smooth_fields(np.random.default_rng(9), 1, 0.3), -1.0, DEX)
    base = part.copy()
    cu, dcu, amu = (np.zeros((d, NY + 1, NX + 1)) for d in (3, 2, 3))
    qpo.amjdp(part, fxy, bxy, psit, cu, dcu, amu, -1.0, DXI, DEX)
    ppart, kpic, _ = ref.to_tiles(base, NX, NY)
    rcu, rdcu, ramu = ref.rfield(3, NX, NY), ref.rfield(2, NX, NY), ref.rfield(3, NX, NY)
    ref.amjdp(ppart, kpic, ref.r_from_oracle(fxy, NX, NY), ref.r_from_oracle(bxy, NX, NY),
              ref.r_from_oracle(psit, NX, NY), rcu, rdcu, ramu, -1.0, DXI, DEX, NX, NY)
    assert np.array_equal(part, ref.from_tiles(ppart, kpic))  # p6 updated in place
    for mine, theirs in ((cu, rcu), (dcu, rdcu), (amu, ramu)):
        assert rel(mine, ref.r_to_oracle(theirs, NX, NY)) < 1e-13


def test_push_and_reorder(ref, plasma):
    fxy, bxy, psit = _fields(8)
    fxy *= 20.0  # strong enough that particles cross tiles and the walls
    part = plasma.copy()
    qpo.extpsi(part, smooth_fields(np.random.default_rng(10), 1, 0.3), -1.0, DEX)
    ppart, kpic, _ = ref.to_tiles(part.copy(), NX, NY)
    dt = 0.9  # large step: O(1 cell) moves, across tile edges and the walls
    err = qpo.push(part, fxy, bxy, psit, -1.0, dt, DEX)
    kp = qpo.order_by_tile(NX, NY, part)
    ppart, kpic, irc = ref.push_and_reorder(
        ppart, kpic, ref.r_from_oracle(fxy, NX, NY), ref.r_from_oracle(bxy, NX, NY),
        ref.r_from_oracle(psit, NX, NY), -1.0, dt, DEX, NX, NY)
    assert err == 0 and irc == 0
    assert np.array_equal(kp, kpic)               # per-tile counts, bit-exact
    rp = ref.from_tiles(ppart, kpic)
    # the same particle multiset, tile by tile (order inside a tile differs)
    for k in range(len(kp)):
        s0, s1 = kp[:k].sum(), kp[:k + 1].sum()
        a = part[s0:s1][np.lexsort(part[s0:s1, ::-1].T)]
        b = rp[s0:s1][np.lexsort(rp[s0:s1, ::-1].T)]
        assert np.array_equal(a, b), k
    # the push really moved particles between tiles and across the walls
    kp0 = np.bincount(qpo.tile_of(NX, NY, plasma) - 1, minlength=len(kp))
    assert not np.array_equal(kp0, kp)
    assert (rp[:, 0] < 1.0).any() and (rp[:, 0] >= NX - 1).any()
    assert (rp[:, 1] < 1.0).any()


def test_ranorm_bit_exact():
    r = fresh_ref()
    want = np.array([r.ranorm() for _ in range(4096)])
    got = qpo.ranorm_stream(4096)
    assert np.array_equal(got, want)


def test_beam_init_bit_exact():
    grid = qpo.Grid(6, 6, 6)
    spec = qpo.BeamSpec(np_=(16, 16, 32), peak=4.0)
    part, qm, ierr = qpo.beam_init(grid, spec)
    r = fresh_ref()
    bp = qpo.beam_params(grid, spec)
    x0 = (bp.bcx / bp.dx, bp.bcy / bp.dy, bp.bcz / bp.dz)
    sig = (bp.sigx / bp.dx, bp.sigy / bp.dy, bp.sigz / bp.dz)
    rpart, rerr = r.beam_dist(qm, grid.nx, grid.ny, grid.nz, 16, 16, 32,
                              spec.sigma_v, (0.0, 0.0, spec.gamma), sig, x0,
                              (0, 0, 0), (0, 0, 0), True, part.shape[0] + 2)
    assert ierr == 0 and rerr == 0
    assert np.array_equal(part, rpart)


def test_beam_deposit_and_push(ref):
    grid = qpo.Grid(5, 5, 5)
    spec = qpo.BeamSpec(np_=(8, 8, 16), peak=4.0, sigma=(1.5, 1.5, 1.5),
                        center=(0.0, 0.0, 5.0))
    part, qm, _ = qpo.beam_init(grid, spec)
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    q3 = np.zeros((1, nz + 1, ny + 1, nx + 1))
    qpo.beam_deposit(part, q3)
    qr = ref.beam_deposit(part, nx, ny, nz)
    assert rel(q3[0], np.transpose(qr[:nx + 1], (2, 1, 0))) < 1e-14
    rng = np.random.default_rng(11)
    e3 = rng.normal(0, 0.2, (3, nz + 1, ny + 1, nx + 1))
    b3 = rng.normal(0, 0.2, (3, nz + 1, ny + 1, nx + 1))
    e3r = np.zeros((3, nx + 2, ny + 1, nz + 1), order="F")
    b3r = np.zeros((3, nx + 2, ny + 1, nz + 1), order="F")
    e3r[:, :nx + 1] = np.transpose(e3, (0, 3, 2, 1))
    b3r[:, :nx + 1] = np.transpose(b3, (0, 3, 2, 1))
    dt = 2000.0  # long enough that some particles leave through every face
    mine = qpo.beam_push(part.copy(), e3, b3, -1.0, dt, grid.dx, grid.dz)
    theirs = ref.beam_push(part.copy(), e3r, b3r, -1.0, dt, grid.dx, grid.dz, nx, ny, nz)
    assert 0 < mine.shape[0] < part.shape[0]
    assert np.array_equal(mine, theirs)  # same survivors, same order


def _composed(indx, indy, species, beams, nslices):
    from oracle import refslice
    grid = qpo.Grid(indx, indy, 5)
    osim = qpo.Simulation(grid, species, beams, dt=10.0, iter_=2)
    rsim = refslice.RefSimulation(fresh_ref(), grid, species, beams, dt=10.0, iter_=2)
    osim.begin_step()
    rsim.begin_step()
    assert np.array_equal(osim.q3[0], np.transpose(rsim.q3[:grid.nx + 1], (2, 1, 0))) or \
        rel(osim.q3[0], np.transpose(rsim.q3[:grid.nx + 1], (2, 1, 0))) < 1e-14
    for j in range(1, nslices + 1):
        osim.slice(j)
        rsim.slice(j)
        for c in range(3):
            # Bz is a small second-order field built from cancelling terms:
            # its bar is relative to its own peak, like the others
            bar = 1e-11 if c == 2 else 1e-12
            assert rel(osim.e3[c, j], rsim.e3[c, j]) < 1e-12, (j, "e3", c)
            assert rel(osim.b3[c, j], rsim.b3[c, j]) < bar, (j, "b3", c)
        for s, r in zip(osim.species, rsim.species):
            assert np.array_equal(s.kpic, r.kpic), j       # per-tile counts, bit-exact
            rp = r.particles()
            # every particle's cell, as a per-cell count (bit-exact)
            ncell = (grid.nx + 1) * (grid.ny + 1)
            assert np.array_equal(
                np.bincount(qpo.cell_of(grid.nx, grid.ny, s.part), minlength=ncell),
                np.bincount(qpo.cell_of(grid.nx, grid.ny, rp), minlength=ncell)), j
            k0 = 0
            for k, n in enumerate(s.kpic):
                a = s.part[k0:k0 + n]
                b = rp[k0:k0 + n]
                # same particles tile by tile (the reference's order inside a
                # tile is its hole-fill order): compare sorted by cell, then x, y
                ka = np.lexsort((a[:, 1], a[:, 0]))
                kb = np.lexsort((b[:, 1], b[:, 0]))
                assert np.abs(a[ka] - b[kb]).max(initial=0.0) < 1e-12, (j, k)
                k0 += n


def test_composed_slices_match_the_reference(ref):
    """qpo.Simulation's slice sequence (simulation_class.f03:643-736) against the
    same loop composed independently over the reference's own kernels
    (oracle/refslice.py): 8 slices of a 64^2 grid through the beam."""
    beam = qpo.BeamSpec(center=(0.0, 0.0, 1.0), sigma=(1.0, 1.0, 0.6), np_=(16, 16, 32),
                        peak=4.0)
    _composed(6, 6, [qpo.SpeciesSpec(ppc=(2, 2))], [beam], 8)


def test_composed_slices_nonsquare_two_species_two_beams(ref):
    """The same on a 64 x 32 grid with an electron and a heavy positive species
    of different ppc and density, and two beams (the per-species additions of
    species2d_class.f03:166-218 and the beam copy-adds of :652-655)."""
    b1 = qpo.BeamSpec(center=(0.0, 0.0, 1.0), sigma=(1.0, 1.0, 0.6), np_=(16, 16, 32), peak=4.0)
    b2 = qpo.BeamSpec(center=(1.0, -0.5, 1.5), sigma=(0.8, 1.2, 0.5), np_=(8, 16, 32), peak=2.0)
    sp = [qpo.SpeciesSpec(ppc=(2, 2)), qpo.SpeciesSpec(ppc=(1, 3), density=0.5, q=1.0, m=100.0)]
    _composed(6, 5, sp, [b1, b2], 8)
