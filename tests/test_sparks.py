"""Spark statistics definitions (CPU only)."""
import json
import math
import os

import numpy as np
import pytest

import csparks as cs
import dyad
import sparks
from conftest import REPO

FIXTURE = os.path.join(REPO, "tests", "golden", "spark_stats_fast_dyad.json")


def test_fdhm_of_a_triangle():
    t = np.linspace(0.0, 10.0, 1001)
    y = np.maximum(0.0, 1.0 - np.abs(t - 4.0) / 2.0)      # half height at t = 3 and 5
    assert sparks.fdhm(t, y) == pytest.approx(2.0, abs=1e-9)


def test_fdhm_of_a_gaussian_interpolates_between_samples():
    t = np.linspace(0.0, 10.0, 41)
    sigma = 1.3
    y = np.exp(-((t - 5.1) ** 2) / (2 * sigma ** 2))
    # closed form 2 sqrt(2 ln 2) sigma; linear interpolation on a 0.25 grid
    assert sparks.fdhm(t, y) == pytest.approx(2 * math.sqrt(2 * math.log(2)) * sigma, rel=2e-2)


def test_fdhm_censors_at_the_end_and_handles_no_signal():
    t = np.linspace(0.0, 1.0, 11)
    assert sparks.fdhm(t, np.linspace(0.0, 1.0, 11)) == pytest.approx(0.5, abs=1e-12)
    assert sparks.fdhm(t, np.zeros(11)) == 0.0


def test_line_through_cleft():
    d = dyad.fast_spark_dyad()
    vox = d.geometry.voxels
    off = sparks.line_through(vox, d.model.domain_names.index("cleft"))
    n = vox.shape
    y, z = n[1] // 2 + off[1], n[2] // 2 + off[2]
    cleft = d.model.domain_names.index("cleft")
    # the line crosses the cleft under the cluster, where the channels are
    assert vox[n[0] // 2, y, z] == cleft
    assert np.sum(vox[:, y, z] == cleft) >= n[0] // 2


def test_fixture_holds_the_oracle_realisations():
    with open(FIXTURE) as f:
        d = json.load(f)
    rows = d["realisations"]
    assert len(rows) >= 1000 and d["rng"] == "mt"
    rec = np.array([r["recruited"] for r in rows])
    # bimodal by construction: the first channel alone, or the whole cluster
    assert 0.2 < np.mean(rec == 1) < 0.5 and 0.5 < np.mean(rec == 16) < 0.8


LATTICE = os.path.join(REPO, "tests", "golden", "spark_stats_lattice_dyad.json")


@pytest.mark.skipif(not os.path.exists(LATTICE), reason="spark_stats_lattice_dyad.json not generated")
def test_lattice_fixture_holds_the_shipped_kinetics_runs():
    """200 oracle realisations of the lattice dyad with the shipped kinetics
    over the first 2 ms: every run starts with the forced-open channel, so
    at least one channel is recruited, and the step plan is the lattice
    dyad's (dt 0.95 h^2 / (6 sigma_max) at 6 nm)."""
    with open(LATTICE) as f:
        d = json.load(f)
    rows = d["realisations"]
    assert len(rows) >= 200 and d["rng"] == "mt" and d["model"] == "dyad.lattice_dyad()"
    assert [r["seed"] for r in rows] == list(range(1, len(rows) + 1))
    dt = 0.95 * 36.0 / (6 * 220000.0)
    assert d["nsteps"] == cs.host_lib().cs_count_steps(d["tstop_ms"], dt)
    assert d["save_every"] == cs.host_lib().cs_save_interval(d["save_ms"], dt)
    rec = np.array([r["recruited"] for r in rows])
    assert rec.min() >= 1
    assert np.all(np.isfinite([r["peak"] for r in rows])) and np.all(np.array([r["duration"] for r in rows]) >= 0)


@pytest.mark.skipif(not os.path.exists(LATTICE), reason="spark_stats_lattice_dyad.json not generated")
def test_lattice_atom_is_the_channel_that_never_closes():
    """The lattice dyad's RyR closes at a constant 0.5/ms, so the forced-open
    channel stays open through the 2 ms window with probability exp(-1);
    with no other channel recruited that run is one deterministic
    trajectory, repeated exactly by every such seed.  Its share of the 200
    oracle runs must match exp(-1) within 4 binomial standard deviations."""
    with open(LATTICE) as f:
        d = json.load(f)
    rows = d["realisations"]
    peak = np.array([r["peak"] for r in rows])
    vals, counts = np.unique(peak, return_counts=True)
    assert counts.max() > 1
    atom = vals[np.argmax(counts)]
    share = float(np.mean(peak == atom))
    p = math.exp(-0.5 * d["tstop_ms"])
    assert abs(share - p) < 4 * math.sqrt(p * (1 - p) / len(rows)), (share, p)
    # the same runs carry the same duration and recruit only the trigger
    dur = np.array([r["duration"] for r in rows])[peak == atom]
    rec = np.array([r["recruited"] for r in rows])[peak == atom]
    assert np.all(dur == dur[0]) and np.all(rec == 1)
    # among the runs that recruit nobody the atom's peak is the largest:
    # its channel never closed
    alone = np.array([r["recruited"] for r in rows]) == 1
    assert atom >= np.max(peak[alone & (peak != atom)])


@pytest.mark.skipif(not os.path.exists(LATTICE), reason="spark_stats_lattice_dyad.json not generated")
def test_ks_comparison_snaps_fp32_copies_of_an_atom():
    """Moving the oracle's atom by 3e-7 (an FP32 rounding of the same
    trajectory) makes the plain two-sample KS test report the atom's mass
    as D and fail; after snapping onto the oracle's atoms the same sample
    passes, while a real shift of the continuous part still fails."""
    from scipy.stats import ks_2samp
    from helpers import snap_to_atoms
    with open(LATTICE) as f:
        rows = json.load(f)["realisations"]
    ref = np.array([r["peak"] for r in rows])
    vals, counts = np.unique(ref, return_counts=True)
    atom = vals[np.argmax(counts)]
    got = ref.copy()
    got[got == atom] *= 1 + 3e-7
    assert ks_2samp(got, ref).pvalue < 1e-6
    snapped, worst, moved, atoms = snap_to_atoms(got, ref)
    assert moved == int(np.sum(ref == atom)) and atom in atoms and worst < 1e-6
    assert ks_2samp(snapped, ref).pvalue > 0.5
    shifted = np.where(ref == atom, ref, ref * 0.8)
    assert ks_2samp(snap_to_atoms(shifted, ref)[0], ref).pvalue < 0.05 / 3
