"""Reads a QuickPIC input deck (qpinput.json) and runs it on the GPU solver.

The deck format is the reference's (input_class.f03:62-145): JSON in which
'!' starts a comment that runs to the end of the line (json-fortran's comment
character), with the keys read in simulation_class.f03:151-224 (grid, box, time
step, corrector passes, counts of beams and species) and 392-569 (per beam and
per species).  The first `nbeams` beams and `nspecies` species are used, as in
the reference.  This build implements beam profiles 0 (bi-Gaussian,
fdist3d_000), 2 (Twiss, fdist3d_002), 3 (particles from a file, fdist3d_003)
and 100 (quiet-start torus, fdist3d_100), and species profiles 0
(uniform, fdist2d_000), 10 (elliptical annulus), 11 (Gaussian ring) and 12
(piecewise radial density), each with a uniform or piecewise longitudinal
density; a deck that asks for another profile is rejected with the
reference's message for it.
Diagnostics (`diag` of fields, species and beams) and beam restart files
(`dump_restart`, `read_restart`) are written and read in the reference's HDF5
format by qpdiag.py.  The MPI layout (`nodes`) has no effect on one GPU.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import qpslice as qps


def strip_comments(text: str) -> str:
    """Removes '!' comments outside strings."""
    out = []
    for line in text.splitlines():
        in_str = False
        esc = False
        cut = len(line)
        for i, ch in enumerate(line):
            if in_str:
                if esc:
                    esc = False
                elif ch == "\\":
                    esc = True
                elif ch == '"':
                    in_str = False
            elif ch == '"':
                in_str = True
            elif ch == "!":
                cut = i
                break
        out.append(line[:cut])
    return "\n".join(out)


def load(path: str) -> dict:
    with open(path) as fh:
        return json.loads(strip_comments(fh.read()))


@dataclass
class BeamCfg:
    np_: tuple
    center: tuple
    sigma: tuple
    sigma_v: tuple
    gamma: float
    peak: float
    q: float = -1.0
    m: float = 1.0
    quiet: bool = True
    evolution: bool = True
    npmax: int = 0
    centroid_x: tuple = (0.0, 0.0, 0.0)
    centroid_y: tuple = (0.0, 0.0, 0.0)


@dataclass
class TwissCfg:
    """Beam profile 2 (init_fdist3d_002, fdist3d_class.f03:652-745)."""
    np_: tuple
    center: tuple
    alpha: tuple
    beta: tuple
    emittance: tuple
    sigma_z: float
    sigma_vz: float
    gamma: float
    peak: float
    q: float = -1.0
    m: float = 1.0
    quiet: bool = True
    evolution: bool = True
    npmax: int = 0


@dataclass
class FileBeamCfg:
    """Beam profile 3 (init_fdist3d_003, fdist3d_class.f03:823-897)."""
    file_name: str
    np_: int
    center: tuple
    n0: float
    q: float = -1.0
    m: float = 1.0
    evolution: bool = True
    npmax: int = 0


@dataclass
class TorusBeamCfg:
    """Beam profile 100 (init_fdist3d_100, fdist3d_class.f03:976-1082)."""
    center: tuple
    range_x: tuple
    range_y: tuple
    range_z: tuple
    ppc: tuple
    q: float
    m: float
    np_: float
    gamma: float
    r1: float
    r2: float
    evolution: bool = True
    npmax: int = 0


@dataclass
class SpeciesCfg:
    ppc: tuple
    density: float = 1.0
    q: float = -1.0
    m: float = 1.0
    piecewise: tuple = None     # (s points, density factors) or None
    profile: int = 0            # 0, 10, 11, 12
    transverse: dict = field(default_factory=dict)   # the profile's keys (set_profile)


@dataclass
class DeckCfg:
    indx: int
    indy: int
    indz: int
    box: tuple            # xmin, xmax, ymin, ymax, zmin, zmax
    dt: float
    time: float
    iter: int
    beams: list = field(default_factory=list)
    species: list = field(default_factory=list)

    @property
    def nstep3d(self) -> int:
        # this%nstep3d = min/max with integer assignment (simulation_class.f03:197-199)
        return int(self.time / self.dt)


def _f3(v):
    return tuple(float(x) for x in v)


def config(deck: dict) -> DeckCfg:
    """The hot-path configuration of a parsed deck (simulation_class.f03:151-224,
    392-569)."""
    s = deck["simulation"]
    if str(s.get("boundary", "conducting")) != "conducting":
        raise ValueError("only conducting boundaries (psolver 1) are supported")
    box = s["box"]
    cfg = DeckCfg(indx=int(s["indx"]), indy=int(s["indy"]), indz=int(s["indz"]),
                  box=(float(box["x"][0]), float(box["x"][1]), float(box["y"][0]),
                       float(box["y"][1]), float(box["z"][0]), float(box["z"][1])),
                  dt=float(s["dt"]), time=float(s["time"]), iter=int(s["iter"]))
    for i in range(int(s["nbeams"])):
        b = deck["beam"][i]
        prof = int(b["profile"])
        if prof == 2:
            cfg.beams.append(TwissCfg(
                np_=tuple(int(x) for x in b["np"]), center=_f3(b["center"]),
                alpha=tuple(float(x) for x in b["alpha"]), beta=tuple(float(x) for x in b["beta"]),
                emittance=tuple(float(x) for x in b["emittance"]),
                sigma_z=float(b["sigmaz"]), sigma_vz=float(b["sigma_vz"]),
                gamma=float(b["gamma"]), peak=float(b["peak_density"]), q=float(b["q"]),
                m=float(b.get("m", 1.0)), quiet=bool(b.get("quiet_start", True)),
                evolution=bool(b.get("evolution", True)), npmax=int(b.get("npmax", 0))))
            continue
        if prof == 3:     # init_fdist3d_003 (fdist3d_class.f03:823-897)
            cfg.beams.append(FileBeamCfg(
                file_name=str(b["file_name"]), np_=int(b["np"]), center=_f3(b["center"]),
                n0=float(s["n0"]), q=float(b["q"]), m=float(b["m"]),
                evolution=bool(b.get("evolution", True)), npmax=int(b.get("npmax", 0))))
            continue
        if prof == 100:   # init_fdist3d_100 (fdist3d_class.f03:976-1082)
            cfg.beams.append(TorusBeamCfg(
                center=_f3(b["center"]), range_x=tuple(float(x) for x in b["range_x"]),
                range_y=tuple(float(x) for x in b["range_y"]),
                range_z=tuple(float(x) for x in b["range_z"]),
                ppc=tuple(int(x) for x in b["ppc"]), q=float(b["q"]), m=float(b["m"]),
                np_=float(b["np"]), gamma=float(b["gamma"]), r1=float(b["r1"]),
                r2=float(b["r2"]), evolution=bool(b.get("evolution", True)),
                npmax=int(b.get("npmax", 0))))
            continue
        if prof != 0:
            raise ValueError(f"Invalid beam profile number: {prof} "
                             "(this build supports 0, 2, 3, 100)")
        cfg.beams.append(BeamCfg(
            np_=tuple(int(x) for x in b["np"]), center=_f3(b["center"]), sigma=_f3(b["sigma"]),
            sigma_v=_f3(b["sigma_v"]), gamma=float(b["gamma"]),
            peak=float(b["peak_density"]), q=float(b["q"]), m=float(b["m"]),
            quiet=bool(b.get("quiet_start", True)), evolution=bool(b.get("evolution", True)),
            npmax=int(b.get("npmax", 0)), centroid_x=_f3(b.get("centroid_x", (0, 0, 0))),
            centroid_y=_f3(b.get("centroid_y", (0, 0, 0)))))
    for i in range(int(s["nspecies"])):
        p = deck["species"][i]
        prof = int(p["profile"])
        if prof not in (0, 10, 11, 12):
            raise ValueError(f"Invalid species profile number: {prof} "
                             "(this build supports 0, 10, 11, 12)")
        tp = {}
        if prof in (10, 11, 12):     # keys of init_fdist2d_010/011/012
            tp["center"] = (float(p["center"][0]), float(p["center"][1]))
        if prof == 10:
            tp["inner_radius"] = (float(p["inner_radius"][0]), float(p["inner_radius"][1]))
            tp["outer_radius"] = (float(p["outer_radius"][0]), float(p["outer_radius"][1]))
        elif prof == 11:
            tp["radius"], tp["width"] = float(p["radius"]), float(p["width"])
        elif prof == 12:
            tp["radial"] = (tuple(float(x) for x in p["piecewise_r"]),
                            tuple(float(x) for x in p["piecewise_radial_density"]))
        lp = p.get("longitudinal_profile", "uniform")
        pw = None
        if lp == "piecewise":     # fdist2d_class.f03:197-201
            pw = (tuple(float(x) for x in p["piecewise_s"]),
                  tuple(float(x) for x in p["piecewise_density"]))
        elif lp != "uniform":
            raise ValueError(f"longitudinal profile '{lp}' is not supported "
                             "(uniform, piecewise)")
        cfg.species.append(SpeciesCfg(ppc=tuple(int(x) for x in p["ppc"]),
                                      density=float(p["density"]), q=float(p["q"]),
                                      m=float(p["m"]), piecewise=pw, profile=prof,
                                      transverse=tp))
    return cfg


class Run:
    """The deck's simulation on one GPU: context, species, beams and the
    slice driver (simulation_class.f03:611-765)."""

    def __init__(self, cfg: DeckCfg, precision=qps.FP64, device=0, seed=0, deck=None,
                 root="."):
        self.cfg = cfg
        self.ctx = qps.Ctx(cfg.indx, cfg.indy, cfg.indz, box=cfg.box, precision=precision,
                           device=device)
        sim_keys = deck["simulation"] if deck else {}
        # start3d (simulation_class.f03:193-201): a restart resumes after its step
        self.start3d = (int(sim_keys["restart_timestep"]) + 1
                        if bool(sim_keys.get("read_restart", False)) else 1)
        s0 = (self.start3d - 1) * cfg.dt
        self.species = [qps.Species(self.ctx, ppc=s.ppc, density=s.density, q=s.q, m=s.m)
                        for s in cfg.species]
        for sp, s in zip(self.species, cfg.species):
            if s.piecewise is not None:
                sp.set_longitudinal(s.piecewise[0], s.piecewise[1], s0)
            if s.profile != 0:
                sp.set_profile(s.profile, s_now=s0, **s.transverse)
        self.beams = []
        for i, b in enumerate(cfg.beams):
            if isinstance(b, FileBeamCfg):
                path = b.file_name if os.path.isabs(b.file_name) else os.path.join(root,
                                                                                    b.file_name)
                self.beams.append(qps.FileBeam(self.ctx, path, b.np_, center=b.center, n0=b.n0,
                                               q=b.q, m=b.m, evolution=b.evolution, dt=cfg.dt,
                                               npmax=b.npmax))
                continue
            if isinstance(b, TorusBeamCfg):
                self.beams.append(qps.TorusBeam(
                    self.ctx, center=b.center, range_x=b.range_x, range_y=b.range_y,
                    range_z=b.range_z, ppc=b.ppc, q=b.q, m=b.m, np_=b.np_, gamma=b.gamma,
                    r1=b.r1, r2=b.r2, evolution=b.evolution, dt=cfg.dt, npmax=b.npmax))
                continue
            if isinstance(b, TwissCfg):
                self.beams.append(qps.TwissBeam(
                    self.ctx, center=b.center, alpha=b.alpha, beta=b.beta, emittance=b.emittance,
                    sigma_z=b.sigma_z, sigma_vz=b.sigma_vz, np_=b.np_, gamma=b.gamma,
                    peak=b.peak, q=b.q, m=b.m, quiet=b.quiet, evolution=b.evolution, dt=cfg.dt,
                    seed=seed + i, npmax=b.npmax))
                continue
            self.beams.append(qps.Beam(self.ctx, center=b.center, sigma=b.sigma,
                                       sigma_v=b.sigma_v, np_=b.np_, gamma=b.gamma,
                                       peak=b.peak, q=b.q, m=b.m, quiet=b.quiet,
                                       evolution=b.evolution, dt=cfg.dt, seed=seed + i,
                                       npmax=b.npmax, centroid_x=b.centroid_x,
                                       centroid_y=b.centroid_y))
        self.sim = qps.Sim(self.ctx, self.species, self.beams, iter_=cfg.iter)
        self.diag = None
        if deck is not None:
            import qpdiag
            if self.start3d > 1:
                qpdiag.read_restart(self, deck, root)
            self.diag = qpdiag.Diagnostics(self, deck, root)

    def step(self, i, tol=0.0, iter_max=None):
        """3-D step i (1-based): every xi slice, the beam push, the
        diagnostics of step i, then the plasma renewal at s = i*dt
        (simulation_class.f03:611-765).  Returns the corrector passes of each
        slice."""
        sim, nz = self.sim, self.ctx.nz
        sim.begin_step()
        passes = [sim.slice(j, iter_max=iter_max, tol=tol)[0] for j in range(1, nz + 1)]
        sim.push_beams()
        if self.diag is not None:
            self.diag.dump(i)
        sim.renew(i * self.cfg.dt)
        return passes

    def run(self, tol=0.0, iter_max=None):
        """go_simulation: the initial dump at step 0 of a fresh run
        (:606-609), then steps start3d..nstep3d."""
        if self.diag is not None and self.start3d == 1:
            self.diag.dump(0)
        return [self.step(i, tol, iter_max)
                for i in range(self.start3d, self.cfg.nstep3d + 1)]

    def close(self):
        self.ctx.close()
