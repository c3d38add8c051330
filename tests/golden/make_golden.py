"""Writes tests/golden/ref_vectors.npz: inputs and the outputs of the
reference's own F77 kernels (oracle/_ref/libqpref.so, compiled from the
reference sources by oracle/Makefile.ref) on small seeded cases.

tests/test_oracle_golden.py checks the oracle against these vectors, so the
oracle stays pinned where the reference tree (and so oracle/_ref) is absent.

Run from the repository root after `make -C oracle -f Makefile.ref`:
    python tests/golden/make_golden.py
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import qpo                    # noqa: E402  (inputs only: loaders, tiling)
from oracle import refharness as rh       # noqa: E402

NX = NY = 32
DEX = 20.0 / NX
DXI = 10.0 / 64


def smooth_fields(rng, dims, amp=0.05):
    y, x = np.mgrid[0:NY + 1, 0:NX + 1]
    out = np.zeros((dims, NY + 1, NX + 1))
    for c in range(dims):
        for _ in range(4):
            kx, ky = rng.integers(1, 5, 2)
            out[c] += amp * rng.normal() * np.sin(np.pi * kx * x / NX + rng.uniform()) \
                * np.cos(np.pi * ky * y / NY + rng.uniform())
    return out


def fresh_ref():
    d = tempfile.mkdtemp()
    p = os.path.join(d, "libqpref_copy.so")
    shutil.copy(rh.REF_SO, p)
    return rh.Ref(p)


def main():
    if not rh.available():
        sys.exit("oracle/_ref/libqpref.so is missing: build it with oracle/Makefile.ref")
    ref = rh.Ref()
    out = {"nx": NX, "ny": NY, "dex": DEX, "dxi": DXI}
    # ---- transforms: every wrapper, both directions ----
    for kind, dim in sorted(rh.Ref.WRAP):
        rng = np.random.default_rng(100 + kind * 10 + dim)
        f = rng.standard_normal((dim, NY + 1, NX + 1))
        g = ref.kfield(dim, NX, NY)
        ref.transform(kind, ref.r_from_oracle(f, NX, NY), g, -1, NX, NY)
        out[f"r2k_{kind}_{dim}_in"] = f
        out[f"r2k_{kind}_{dim}_out"] = ref.k_to_oracle(g)
        k = rng.standard_normal((dim, NX + 1, NY + 1))
        fr = ref.rfield(dim, NX, NY)
        ref.transform(kind, fr, ref.k_from_oracle(k), 1, NX, NY)
        ref.copyguard(fr, NY)
        out[f"k2r_{kind}_{dim}_in"] = k
        out[f"k2r_{kind}_{dim}_out"] = ref.r_to_oracle(fr, NX, NY)
    # ---- Poisson table and k-space operators ----
    ffc = ref.pois_table(NX, NY)
    out["pois_g"], out["pois_s"] = ffc.real.T.copy(), ffc.imag.T.copy()
    rng = np.random.default_rng(3)
    ks = lambda d: rng.standard_normal((d, NX + 1, NY + 1))   # noqa: E731
    q = ks(1)
    out["kop_q"] = q
    for name, dout in (("pot", 1), ("smooth", 1), ("elf", 2)):
        o = ref.kfield(dout, NX, NY)
        qin = ref.k_from_oracle(q)
        if dout == 1:
            o1 = np.asfortranarray(o[0])
            ref.kcall(name, ffc, NX, NY, np.asfortranarray(qin[0]), o1)
            o[0] = o1
        else:
            ref.kcall(name, ffc, NX, NY, np.asfortranarray(qin[0]), o)
        out[f"kop_{name}"] = ref.k_to_oracle(o)
    cu = ks(3)
    o = ref.kfield(3, NX, NY)
    ref.kcall("bf", ffc, NX, NY, ref.k_from_oracle(cu), o)
    out["kop_cu"], out["kop_bf"] = cu, ref.k_to_oracle(o)
    f2 = ks(2)
    o = ref.kfield(1, NX, NY)
    o1 = np.asfortranarray(o[0])
    ref.kcall("div", ffc, NX, NY, ref.k_from_oracle(f2), o1)
    out["kop_f2"], out["kop_div"] = f2, ref.k_to_oracle(o1[None])
    o = ref.kfield(2, NX, NY)
    ref.kcall("grad", ffc, NX, NY, np.asfortranarray(ref.k_from_oracle(q)[0]), o)
    out["kop_grad"] = ref.k_to_oracle(o)
    # ---- B-perp corrector ----
    cu, dcu, amu, b = ks(3), ks(2), ks(3), ks(2)
    out["bfqp_in"] = [cu.copy(), dcu.copy(), amu.copy(), b.copy()]
    rcu, rdcu, ramu, rb = (ref.k_from_oracle(a) for a in (cu, dcu, amu, b))
    ref.bfqp(ffc, rcu, rdcu, ramu, rb, DEX * DEX, DEX, NX, NY)
    out["bfqp_b"], out["bfqp_dcu"] = ref.k_to_oracle(rb), ref.k_to_oracle(rdcu)
    # ---- plasma: loader, charge, psi gather, current deposit, push ----
    part = qpo.plasma_init(NX, NY, 2, 2)
    qpo.order_by_tile(NX, NY, part)
    rng = np.random.default_rng(7)
    part[:, 2:5] = rng.normal(0.0, 0.3, (part.shape[0], 3))
    out["plasma"] = part.copy()
    ppart, kpic, _ = ref.to_tiles(part.copy(), NX, NY)
    qr = ref.rfield(1, NX, NY)
    ref.qdeposit(ppart, kpic, qr, NX, NY)
    ref.fold(qr, NY)
    out["qdeposit"] = ref.r_to_oracle(qr, NX, NY)
    rng = np.random.default_rng(5)
    psi = smooth_fields(rng, 1, amp=0.5)
    ppart, kpic, _ = ref.to_tiles(part.copy(), NX, NY)
    ref.extpsi(ppart, kpic, ref.r_from_oracle(psi, NX, NY), -1.0, DEX, NX, NY)
    out["extpsi_psi"], out["extpsi_out"] = psi, ref.from_tiles(ppart, kpic)
    rng = np.random.default_rng(6)
    fxy, bxy, psit = smooth_fields(rng, 2), smooth_fields(rng, 3), smooth_fields(rng, 1)
    base = out["extpsi_out"].copy()
    ppart, kpic, _ = ref.to_tiles(base.copy(), NX, NY)
    rcu, rdcu, ramu = ref.rfield(3, NX, NY), ref.rfield(2, NX, NY), ref.rfield(3, NX, NY)
    ref.amjdp(ppart, kpic, ref.r_from_oracle(fxy, NX, NY), ref.r_from_oracle(bxy, NX, NY),
              ref.r_from_oracle(psit, NX, NY), rcu, rdcu, ramu, -1.0, DXI, DEX, NX, NY)
    out["amjdp_f"] = [fxy, bxy, psit]   # particles in: extpsi_out
    out["amjdp_part"] = ref.from_tiles(ppart, kpic)
    out["amjdp_cu"] = ref.r_to_oracle(rcu, NX, NY)
    out["amjdp_dcu"] = ref.r_to_oracle(rdcu, NX, NY)
    out["amjdp_amu"] = ref.r_to_oracle(ramu, NX, NY)
    fxy20 = fxy * 20.0
    dt = 0.9
    ppart, kpic, _ = ref.to_tiles(base.copy(), NX, NY)
    ppart, kpic, irc = ref.push_and_reorder(
        ppart, kpic, ref.r_from_oracle(fxy20, NX, NY), ref.r_from_oracle(bxy, NX, NY),
        ref.r_from_oracle(psit, NX, NY), -1.0, dt, DEX, NX, NY)
    assert irc == 0
    out["push_dt"], out["push_fxy"] = dt, fxy20
    out["push_kpic"] = np.asarray(kpic).copy()
    out["push_part"] = ref.from_tiles(ppart, kpic)
    # ---- beam: ranorm stream, loader, deposit, push ----
    r = fresh_ref()
    out["ranorm"] = np.array([r.ranorm() for _ in range(4096)])
    grid = qpo.Grid(5, 5, 5)
    spec = qpo.BeamSpec(np_=(8, 8, 16), peak=4.0, sigma=(1.5, 1.5, 1.5), center=(0.0, 0.0, 5.0))
    _, qm, _ = qpo.beam_init(grid, spec)   # qm only; the particles come from the reference
    r = fresh_ref()
    bp = qpo.beam_params(grid, spec)
    x0 = (bp.bcx / bp.dx, bp.bcy / bp.dy, bp.bcz / bp.dz)
    sig = (bp.sigx / bp.dx, bp.sigy / bp.dy, bp.sigz / bp.dz)
    npt = 8 * 8 * 16
    bpart, rerr = r.beam_dist(qm, grid.nx, grid.ny, grid.nz, 8, 8, 16, spec.sigma_v,
                              (0.0, 0.0, spec.gamma), sig, x0, (0, 0, 0), (0, 0, 0), True,
                              npt + 2)
    assert rerr == 0
    out["beam_part"], out["beam_qm"] = bpart, qm
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    qr = ref.beam_deposit(bpart, nx, ny, nz)
    out["beam_q3"] = np.transpose(qr[:nx + 1], (2, 1, 0)).copy()
    rng = np.random.default_rng(11)
    # inputs rounded to float32 so the fixture can store them at half size
    e3 = rng.normal(0, 0.2, (3, nz + 1, ny + 1, nx + 1)).astype(np.float32).astype(np.float64)
    b3 = rng.normal(0, 0.2, (3, nz + 1, ny + 1, nx + 1)).astype(np.float32).astype(np.float64)
    e3r = np.zeros((3, nx + 2, ny + 1, nz + 1), order="F")
    b3r = np.zeros((3, nx + 2, ny + 1, nz + 1), order="F")
    e3r[:, :nx + 1] = np.transpose(e3, (0, 3, 2, 1))
    b3r[:, :nx + 1] = np.transpose(b3, (0, 3, 2, 1))
    bdt = 2000.0
    out["beam_e3"], out["beam_b3"], out["beam_dt"] = e3.astype(np.float32), b3.astype(np.float32), bdt
    out["beam_push"] = ref.beam_push(bpart.copy(), e3r, b3r, -1.0, bdt, grid.dx, grid.dz,
                                     nx, ny, nz)
    flat = {}
    for k, v in out.items():
        if isinstance(v, list):
            for i, a in enumerate(v):
                flat[f"{k}_{i}"] = np.asarray(a)
        else:
            flat[k] = np.asarray(v)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **flat)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB, {len(flat)} arrays)")


if __name__ == "__main__":
    main()
