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
    deck["simulation"]["nbeams"] = 2          # the second beam has profile 3
    with pytest.raises(ValueError, match="Invalid beam profile number: 3"):
        qpdeck.config(deck)
    deck = qpdeck.load(DECK)
    deck["species"][0]["longitudinal_profile"] = "gaussian"
    with pytest.raises(ValueError, match="gaussian"):
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
