"""File formats, problem setup on the reference's own inputs, and the
command line (CPU only: everything here stops before the GPU is used)."""
import os
import subprocess

import numpy as np
import pytest

import csparks as cs
import dyad
import oracle_py as op
from helpers import load_geometry_fixture, load_model_fixture, with_serca

CLI = os.path.join(cs.BUILD, "calcium_sparks")


def test_g1_fixture_matches_reference_counts():
    """G1_nonP_0.h5 as shipped: 84^3 voxels at 12 nm, 31416 SERCA pairs,
    50 RyRs, 708 sr_cleft pairs; cyt/cleft and jsr/nsr connected."""
    g = load_geometry_fixture("g1_nonP_0")
    assert g.voxels.shape == (84, 84, 84) and g.h == 12.0
    assert [(b.name, b.type, len(b.pairs)) for b in g.boundaries] == [
        ("serca", "density", 31416), ("ryr", "discrete", 50), ("sr_cleft", "density", 708)]
    assert list(np.bincount(g.voxels.ravel(), minlength=5)) == [580787, 708, 2041, 7752, 1416]
    assert g.domain_connections[0, 1] == g.domain_connections[2, 3] == 1
    assert g.domain_connections[0, 2] == 0


def test_geometry_species_roundtrip(tmp_path):
    d = with_serca(dyad.lattice_dyad(seed=3))
    gp, mp = str(tmp_path / "g.h5"), str(tmp_path / "m.h5")
    cs.write_geometry(gp, d.geometry)
    cs.write_species(mp, d.model)
    g2, m2 = cs.read_geometry(gp), cs.read_species(mp)
    np.testing.assert_array_equal(g2.voxels, d.geometry.voxels)
    for a, b in zip(g2.boundaries, d.geometry.boundaries):
        assert (a.name, a.type) == (b.name, b.type)
        np.testing.assert_array_equal(a.pairs, b.pairs)
    assert m2.species_names == d.model.species_names
    np.testing.assert_array_equal(m2.init, d.model.init)
    assert m2.ryr == d.model.ryr and m2.p_ryr == d.model.p_ryr
    assert m2.buffers == d.model.buffers and m2.sigma == d.model.sigma


def test_g1_substeps_match_reference_run():
    """The reference run on G1 at 12 nm with -d 1e-4 reported substeps
    Ca 1, Fluo 11, CMDN 10, ATP 1 (its -v time-stepping table)."""
    g = load_geometry_fixture("g1_nonP_0")
    m = load_model_fixture("params_sr1100_ryr55")
    s = cs.Setup(g, m, cs.Options(h=12.0, dt=1e-4, seed=1, open={"ryr": [24]}, t_close={"ryr": 5.0}))
    assert list(s.problem.substeps)[:5] == [1, 11, 10, 1, 0]
    assert s.problem.dt == 1e-4


@pytest.mark.parametrize("h", [12.0, 6.0])
def test_g1_setup_matches_oracle(h):
    g = load_geometry_fixture("g1_nonP_0")
    m = load_model_fixture("params_sr1100_ryr55")
    opts = cs.Options(h=h, seed=2)
    s = cs.Setup(g, m, opts)
    o = op.Oracle(g, m, opts)
    assert s.n == o.n == tuple(int(84 * 12 / h) for _ in range(3))
    assert s.problem.dt == o.dt
    np.testing.assert_array_equal(np.array(s.problem.coef), np.array(o.info.coef))
    serca = [s.problem.fluxes[f] for f in range(s.problem.num_fluxes) if s.problem.fluxes[f].kind == cs.FLUX_SERCA][0]
    assert serca.num_pairs == 31416 * (12 / h) ** 2
    np.testing.assert_array_equal(
        np.array([s.problem.fluxes[0].init_states[i] for i in range(50)]), o.states(0))
    o.close()


def test_pryr_file_needs_pryr_boundary():
    """The phosphorylation species file names a p_ryr boundary that G1 lacks:
    setup fails like the reference's Geometry_get_boundary_id."""
    g = load_geometry_fixture("g1_nonP_0")
    m = load_model_fixture("params_sr900_ryr55_pryr25")
    with pytest.raises(cs.CSError, match="p_ryr"):
        cs.Setup(g, m, cs.Options(seed=1))


def _write_inputs(tmp_path, d):
    gp, mp = str(tmp_path / "geom.h5"), str(tmp_path / "species.h5")
    cs.write_geometry(gp, d.geometry)
    cs.write_species(mp, d.model)
    return gp, mp


def _cli(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)


def test_cli_requires_geometry(tmp_path):
    r = _cli(["-m", "x.h5"], tmp_path)
    assert r.returncode == 1 and "Expected a geometry file" in r.stderr


def test_cli_rejects_open_index_beyond_boundary(tmp_path):
    d = dyad.lattice_dyad()
    gp, mp = _write_inputs(tmp_path, d)
    r = _cli(["-g", gp, "-m", mp, "-q", "-O", "ryr", "3", "60"], tmp_path)
    assert r.returncode == 1
    assert "Expected max ryr indices" in r.stderr and "60" in r.stderr


def test_cli_rejects_non_increasing_open_indices(tmp_path):
    d = dyad.lattice_dyad()
    gp, mp = _write_inputs(tmp_path, d)
    r = _cli(["-g", gp, "-m", mp, "-O", "ryr", "5", "2"], tmp_path)
    assert r.returncode == 1 and "monotonicly" in r.stderr


def test_cli_rejects_unknown_save_species(tmp_path):
    d = dyad.lattice_dyad()
    gp, mp = _write_inputs(tmp_path, d)
    r = _cli(["-g", gp, "-m", mp, "-q", "-s", "Nope"], tmp_path)
    assert r.returncode == 1 and '"Nope" cannot be found in' in r.stderr


def test_cli_rejects_mpi_layout_of_several_ranks(tmp_path):
    r = _cli(["--mpi_dist", "2", "2", "1", "-g", "a", "-m", "b"], tmp_path)
    assert r.returncode == 1 and "--mpi_dist" in r.stderr


def test_cli_prints_arguments_table(tmp_path):
    """Without -q the reference echoes its arguments first (options.c:395-500)."""
    r = _cli(["-m", "m.h5", "-t", "2", "-C", "ryr", "0.5", "-O", "ryr", "1", "3", "-S", "4"], tmp_path)
    assert r.returncode == 1  # no geometry file
    r = _cli(["-g", "missing.h5", "-m", "m.h5", "-t", "2", "-C", "ryr", "0.5", "-O", "ryr", "1", "3"], tmp_path)
    assert "tstop:             2.000000 ms" in r.stdout
    assert "t_close:           ryr 0.5" in r.stdout
    assert "open states:       ryr: 1, 3" in r.stdout
    assert r.returncode == 1 and "cannot open geometry file" in r.stderr
