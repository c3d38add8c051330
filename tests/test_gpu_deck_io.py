"""A deck run on the GPU with the reference's diagnostics and restart keys
(qpdeck + qpdiag, simulation_class.f03:771-1352): the HDF5 files it writes,
read back with HDF5's own h5dump, against the oracle's step
(oracle/qpo.py Simulation with its 3-D diagnostics) at 1e-10 relative; a run
restarted from the restart file written after step 1 gives bitwise the
files of the uninterrupted run at step 2; a piecewise longitudinal density
matches the oracle's.
"""
import copy
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import qpo

qps = pytest.importorskip("qpslice")
qpdeck = pytest.importorskip("qpdeck")

pytestmark = pytest.mark.gpu

H5DUMP = "/opt/conda/bin/h5dump"
DECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "deck_small.json")


def read_h5(path, dset, tmp):
    out = os.path.join(tmp, "x.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path], capture_output=True,
                   check=True)
    return np.fromfile(out, dtype="<f8")


def rel(a, b):
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300))


def deck_with_diags():
    d = qpdeck.load(DECK)
    d["simulation"]["time"] = 20.1                # two 3-D steps
    d["simulation"]["dump_restart"] = True
    d["simulation"]["ndump_restart"] = 1
    d["field"] = {"diag": [{"name": ["ex", "ey", "ez", "bx", "by", "bz", "psi"], "ndump": 1},
                           {"name": ["ez", "ex"], "ndump": 1,
                            "slice": [["xz", 17], ["yz", 9], ["xy", 20]]}]}
    d["species"][0]["diag"] = [{"name": ["charge", "jx", "jy", "jz"], "ndump": 1}]
    d["beam"][0]["diag"] = [{"name": ["charge"], "ndump": 1},
                            {"name": ["raw"], "ndump": 1, "sample": 3}]
    return d


def oracle_run(cfg, steps):
    """The oracle's steps, capturing what each dump reads (after the beam
    push, before the renewal)."""
    grid = qpo.Grid(cfg.indx, cfg.indy, cfg.indz,
                    box=((cfg.box[0], cfg.box[1]), (cfg.box[2], cfg.box[3]),
                         (cfg.box[4], cfg.box[5])))
    b = cfg.beams[0]
    beams = [qpo.BeamSpec(center=b.center, sigma=b.sigma, sigma_v=b.sigma_v, np_=b.np_,
                          gamma=b.gamma, peak=b.peak, q=b.q, m=b.m)]
    sp = [qpo.SpeciesSpec(ppc=s.ppc, density=s.density, q=s.q, m=s.m, piecewise=s.piecewise)
          for s in cfg.species]
    osim = qpo.Simulation(grid, sp, beams, dt=cfg.dt, iter_=cfg.iter, keep_q3=True,
                          keep_diag=True)
    out = {}
    for i in range(1, steps + 1):
        osim.begin_step()
        for j in range(1, grid.nz + 1):
            osim.slice(j)
        osim.push_beams()
        out[i] = {"e3": osim.e3.copy(), "b3": osim.b3.copy(), "psi3": osim.psi3.copy(),
                  "cu3": osim.cu3.copy(), "q3s": osim.q3s[0].copy(), "q3": osim.q3[0].copy(),
                  "beam": osim.beams[0][0].copy()}
        for s in osim.species:
            s.renew(i * cfg.dt)
        osim.q3_initial()
    return out


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_deck_diagnostics_and_restart(tmp_path):
    print("[gpu-test] deck diagnostics and restart", flush=True)
    deck = deck_with_diags()
    cfg = qpdeck.config(deck)
    root = str(tmp_path / "run")
    run = qpdeck.Run(cfg, precision=qps.FP64, deck=deck, root=root)
    try:
        run.run()
    finally:
        run.close()
    want = oracle_run(cfg, 2)
    nx, ny, nz = 2 ** cfg.indx, 2 ** cfg.indy, 2 ** cfg.indz
    tmp = str(tmp_path)
    f = lambda rel_, ds, n: read_h5(os.path.join(root, rel_.format(n=n)), ds, tmp)  # noqa: E731
    for n in (1, 2):
        w = want[n]
        grids = {
            ("Fields/Ex/ex_{n:08d}.h5", "/ex"): w["e3"][0],
            ("Fields/Ey/ey_{n:08d}.h5", "/ey"): w["e3"][1],
            ("Fields/Ez/ez_{n:08d}.h5", "/ez"): w["e3"][2],
            ("Fields/Bx/bx_{n:08d}.h5", "/bx"): w["b3"][0],
            ("Fields/By/by_{n:08d}.h5", "/by"): w["b3"][1],
            ("Fields/Bz/bz_{n:08d}.h5", "/bz"): w["b3"][2],
            ("Fields/Psi/psi_{n:08d}.h5", "/psi"): w["psi3"],
            ("Species0001/Charge/charge_{n:08d}.h5", "/charge"): w["q3s"],
            ("Species0001/Jx/jx_{n:08d}.h5", "/jx"): w["cu3"][0],
            ("Species0001/Jz/jz_{n:08d}.h5", "/jz"): w["cu3"][2],
            ("Beam0001/Charge/charge_{n:08d}.h5", "/charge"): w["q3"],
        }
        for (path, ds), arr in grids.items():
            got = f(path, ds, n)
            exp = arr[0:nz, 0:ny, 0:nx].ravel()      # planes lpos 1..nz, nodes 0..n-1
            assert got.shape == exp.shape, path
            assert rel(got, exp) < 1e-10, (path, rel(got, exp))
        # slices: xz at y node 16, yz at x node 8, xy at plane 20
        assert rel(f("Fields/Ez_slice0001/ezslicexz_{n:08d}.h5", "/ezslicexz", n),
                   w["e3"][2][0:nz, 16, 0:nx].ravel()) < 1e-10
        assert rel(f("Fields/Ez_slice0002/ezsliceyz_{n:08d}.h5", "/ezsliceyz", n),
                   w["e3"][2][0:nz, 0:ny, 8].ravel()) < 1e-10
        assert rel(f("Fields/Ex_slice0003/exslicexy_{n:08d}.h5", "/exslicexy", n),
                   w["e3"][0][19, 0:ny, 0:nx].ravel()) < 1e-10
        # the raw beam (every 3rd particle) and the restart file, after the push
        key = lambda a: a[np.lexsort(a[:, :3].T[::-1])]   # noqa: E731
        rst = f("RST/Beam-0001/RST-beam0001-0000000000_{n:08d}.h5",
                "/RST-beam0001-0000000000", n).reshape(-1, 7)
        assert rel(key(rst), key(w["beam"])) < 1e-10
        x1 = f("Beam0001/Raw/raw_{n:08d}.h5", "/x1", n)
        assert x1.shape[0] == w["beam"].shape[0] // 3
        assert rel(x1, rst[0:3 * x1.shape[0]:3, 0] * (20.0 / nx) - 10.0) < 1e-15
    # step 0 (the initial dump): zero wake, the plasma charge's first plane
    assert not f("Fields/Ez/ez_{n:08d}.h5", "/ez", 0).any()
    q0 = f("Species0001/Charge/charge_{n:08d}.h5", "/charge", 0).reshape(nz, ny, nx)
    assert rel(q0[0], want[1]["q3s"][0, 0:ny, 0:nx]) < 1e-10 and not q0[1:].any()
    # restart from the file of step 1: step 2 as in the uninterrupted run
    deck2 = copy.deepcopy(deck)
    deck2["simulation"]["read_restart"] = True
    deck2["simulation"]["restart_timestep"] = 1
    root2 = str(tmp_path / "restarted")
    import shutil
    shutil.copytree(os.path.join(root, "RST"), os.path.join(root2, "RST"))
    run2 = qpdeck.Run(qpdeck.config(deck2), precision=qps.FP64, deck=deck2, root=root2)
    try:
        assert run2.start3d == 2
        run2.run()
    finally:
        run2.close()
    for path, ds in (("Fields/Ez/ez_{n:08d}.h5", "/ez"), ("Fields/By/by_{n:08d}.h5", "/by"),
                     ("Beam0001/Charge/charge_{n:08d}.h5", "/charge")):
        a = read_h5(os.path.join(root, path.format(n=2)), ds, tmp)
        b = read_h5(os.path.join(root2, path.format(n=2)), ds, tmp)
        assert np.array_equal(a, b), path
    assert not os.path.exists(os.path.join(root2, "Fields/Ez/ez_00000000.h5"))


def test_piecewise_longitudinal_density_matches_oracle():
    """fdist2d_000 with a piecewise density: the plasma of step 2 (renewed at
    s = dt) and its slices against the oracle at 1e-10."""
    print("[gpu-test] piecewise density", flush=True)
    d = qpdeck.load(DECK)
    d["simulation"]["time"] = 20.1
    d["species"][0]["longitudinal_profile"] = "piecewise"
    d["species"][0]["piecewise_s"] = [0.0, 5.0, 30.0]
    d["species"][0]["piecewise_density"] = [1.0, 0.5, 2.0]
    cfg = qpdeck.config(d)
    assert cfg.species[0].piecewise == ((0.0, 5.0, 30.0), (1.0, 0.5, 2.0))
    run = qpdeck.Run(cfg, precision=qps.FP64)      # loads the plasma at s = 0
    try:
        run.step(1)
        part = run.species[0].download()
        # renewed at s = dt = 10: factor 0.5 + 1.5 * 5 / 25 = 0.8
        assert np.allclose(part[:, 7], -0.8 / 4, rtol=1e-15)
        run.step(2)
        e3 = run.sim.bexyz.download()
    finally:
        run.close()
    want = oracle_run(cfg, 2)[2]["e3"]
    assert rel(e3, want) < 1e-10
