"""A converted phosphorylation pattern on the GPU: G1 with the 'inner 20%'
mask (40 RyRs, 10 pRyRs) and the reference's phosphorylation species file
(parameters_sr900_ryr55_pryr25.h5, whose p_ryr model needs the 'p_ryr'
boundary the converter adds), deterministic channels, FP64, against the
oracle at 1e-10 per voxel."""
import json
import os
import sys

import numpy as np
import pytest

import csparks as cs
import phospho
from helpers import GOLDEN, load_geometry_fixture, load_model_fixture, species_close
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def test_inner_pattern_on_g1_matches_the_oracle():
    m = json.load(open(os.path.join(GOLDEN, "g1_phospho_masks.json")))["G1_inner_10"]
    g = phospho.pattern_geometry(load_geometry_fixture("g1_nonP_0"),
                                 phospho.mask_from_pixels(m["shape"], m["ryr"], m["p_ryr"]))
    model = load_model_fixture("params_sr900_ryr55_pryr25")
    opts = cs.Options(h=12.0, seed=3, open={"ryr": [5, 20], "p_ryr": [0, 7]}, t_close={"ryr": 0.05, "p_ryr": 0.03})
    print("case: G1 inner_10 + sr900_ryr55_pryr25, FP64, 240 steps", flush=True)
    sim = cs.Simulation(g, model, opts)
    orc = Oracle(g, model, opts)
    for _ in range(2):
        sim.step(120)
        orc.step(120)
    sim.synchronize()
    worst = 0.0
    for s in range(orc.ns):
        ok, err = species_close(sim.species(s), orc.species(s), 1e-10)
        worst = max(worst, err)
        assert ok, f"species {s}: {err:.3e}"
    for k in range(2):
        np.testing.assert_array_equal(sim.channel_states(k), orc.states(k))
    print(f"max rel err {worst:.2e}", flush=True)
    sys.stdout.flush()
    sim.close()
    orc.close()
