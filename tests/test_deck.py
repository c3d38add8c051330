"""The input-deck reader (qpinput.json, input_class.f03:62-145) and a deck run
on the GPU against the oracle's slice loop on the same configuration."""
import json
import os

import numpy as np
import pytest

qpdeck = pytest.importorskip("qpdeck")

DECK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "deck_small.json")


def test_comments_are_stripped_outside_strings():
    text = '{"a" : 1, !comment\n "b" : "x!y" ! more\n}'
    assert json.loads(qpdeck.strip_comments(text)) == {"a": 1, "b": "x!y"}


def test_deck_config():
    cfg = qpdeck.config(qpdeck.load(DECK))
    assert (cfg.indx, cfg.indy, cfg.indz) == (5, 5, 5)
    assert cfg.box == (-10.0, 10.0, -10.0, 10.0, 0.0, 10.0)
    assert cfg.nstep3d == 1 and cfg.iter == 2
    assert len(cfg.beams) == 1 and len(cfg.species) == 1     # nbeams, nspecies
    b = cfg.beams[0]
    assert b.np_ == (16, 16, 32) and b.peak == 4.0 and b.sigma == (1.0, 1.0, 0.6)
    assert cfg.species[0].ppc == (2, 2)


def test_unsupported_profile_is_rejected():
    deck = qpdeck.load(DECK)
    deck["simulation"]["nbeams"] = 2
    deck["beam"][1]["profile"] = 1            # RANDOM_NUMBER's stream: not built
    with pytest.raises(ValueError, match="Invalid beam profile number: 1"):
        qpdeck.config(deck)
    deck = qpdeck.load(DECK)
    deck["species"][0]["longitudinal_profile"] = "gaussian"
    with pytest.raises(ValueError, match="gaussian"):
        qpdeck.config(deck)


def test_file_and_torus_beams_are_read():
    """init_fdist3d_003 and init_fdist3d_100 keys (fdist3d_class.f03:
    823-1082); profile 3 takes n0 from the simulation block."""
    deck = qpdeck.load(DECK)
    deck["simulation"]["nbeams"] = 2
    deck["beam"][1].update({"file_name": "beam.txt", "center": [0.0, 0.0, 1.0],
                            "evolution": False})
    cfg = qpdeck.config(deck)
    b = cfg.beams[1]
    assert isinstance(b, qpdeck.FileBeamCfg)
    assert (b.file_name, b.np_, b.n0, b.evolution) == ("beam.txt", 1000, 1e16, False)
    deck["beam"][1] = {"profile": 100, "center": [0.0, 0.0, 2.0], "range_x": [-3, 3],
                       "range_y": [-2, 2], "range_z": [1, 3], "ppc": [2, 2, 1], "q": -1.0,
                       "m": 1.0, "np": 0.5, "gamma": 100.0, "r1": 1.0, "r2": 0.3}
    b = qpdeck.config(deck).beams[1]
    assert isinstance(b, qpdeck.TorusBeamCfg)
    assert b.ppc == (2, 2, 1) and b.range_y == (-2.0, 2.0) and b.r2 == 0.3


def test_transverse_plasma_profiles_are_read():
    """init_fdist2d_010/011/012 keys (fdist2d_class.f03:318-776)."""
    deck = qpdeck.load(DECK)
    sp = deck["species"][0]
    sp.update({"profile": 11, "center": [0.5, -0.5], "radius": 3.0, "width": 0.7})
    cfg = qpdeck.config(deck)
    assert cfg.species[0].profile == 11
    assert cfg.species[0].transverse == {"center": (0.5, -0.5), "radius": 3.0, "width": 0.7}
    sp.update({"profile": 12, "piecewise_r": [0.0, 2.0, 5.0],
               "piecewise_radial_density": [1.0, 0.5, 0.0]})
    cfg = qpdeck.config(deck)
    assert cfg.species[0].transverse["radial"] == ((0.0, 2.0, 5.0), (1.0, 0.5, 0.0))
    sp.update({"profile": 10, "inner_radius": [1.0, 2.0], "outer_radius": [3.0, 4.0]})
    cfg = qpdeck.config(deck)
    assert cfg.species[0].transverse["outer_radius"] == (3.0, 4.0)
    sp["profile"] = 13
    with pytest.raises(ValueError, match="Invalid species profile number: 13"):
        qpdeck.config(deck)


def test_piecewise_longitudinal_profile_is_read():
    """fdist2d_class.f03:197-201: piecewise_s and piecewise_density."""
    deck = qpdeck.load(DECK)
    sp = deck["species"][0]
    sp["longitudinal_profile"] = "piecewise"
    sp["piecewise_s"] = [0, 5, 30]
    sp["piecewise_density"] = [1, 0.5, 2]
    cfg = qpdeck.config(deck)
    assert cfg.species[0].piecewise == ((0.0, 5.0, 30.0), (1.0, 0.5, 2.0))
    del sp["piecewise_s"]
    with pytest.raises(KeyError):
        qpdeck.config(deck)


def test_twiss_beam_is_read():
    """A profile-2 beam (init_fdist3d_002 keys) becomes a Twiss beam."""
    deck = qpdeck.load(DECK)
    deck["beam"][0].update({"profile": 2, "alpha": [0.7, -0.4], "beta": [10.0, 8.0],
                            "emittance": [3.9, 5.0], "sigmaz": 0.6, "sigma_vz": 3.0})
    cfg = qpdeck.config(deck)
    b = cfg.beams[0]
    assert isinstance(b, qpdeck.TwissCfg)
    assert b.alpha == (0.7, -0.4) and b.beta == (10.0, 8.0) and b.emittance == (3.9, 5.0)
    assert b.sigma_z == 0.6 and b.sigma_vz == 3.0 and b.np_ == (16, 16, 32)


@pytest.mark.gpu
def test_deck_run_matches_oracle():
    from oracle import qpo
    import qpslice as qps
    print("[gpu-test] deck run", flush=True)
    cfg = qpdeck.config(qpdeck.load(DECK))
    b = cfg.beams[0]
    osim = qpo.Simulation(qpo.Grid(cfg.indx, cfg.indy, cfg.indz),
                          [qpo.SpeciesSpec(ppc=cfg.species[0].ppc)],
                          [qpo.BeamSpec(center=b.center, sigma=b.sigma, sigma_v=b.sigma_v,
                                        np_=b.np_, gamma=b.gamma, peak=b.peak)],
                          dt=cfg.dt, iter_=cfg.iter)
    osim.begin_step()
    run = qpdeck.Run(cfg, precision=qps.FP64)
    try:
        run.sim.begin_step()
        for j in range(1, 5):
            osim.slice(j)
            run.sim.slice(j)
        got = run.sim.field(qps.F_PSIT).download(0)
        assert np.abs(got - osim.psit).max() / np.abs(osim.psit).max() < 1e-10
    finally:
        run.close()
