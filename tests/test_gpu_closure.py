"""The learned closure on the GPU: schedule replay between a 3 nm and a 12
nm ensemble of the same 12 nm dyads, the closure's inputs against the
fields they are read from, training on paired samples, what the trained
network gains over the plain coarse grid on held-out dyads, and gating on
its predictions inside an ensemble (with the range monitor)."""
import sys

import numpy as np
import pytest

import closure
import csparks as cs
import dyad
import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dyads(seeds, edge=(192.0, 256.0), ryr=(9, 30)):
    return [dyad.ensemble_dyad(s, h=12.0, edge_nm=edge, ryr=ryr) for s in seeds]


def test_replay_copies_schedules_and_features_read_the_fields():
    ds = _dyads(range(4))
    mk = lambda h: cs.Ensemble([(d.geometry, closure.schedule_model(d.model),
                                 cs.Options(h=h, seed=10 + k, open={"ryr": [d.open_index]}))
                                for k, d in enumerate(ds)])
    fine, coarse = mk(3.0), mk(12.0)
    coarse.set_gating(cs.Ensemble.GATE_REPLAY)
    ratio = int(round(coarse.dt / fine.dt))
    assert ratio == 16
    for _ in range(5):
        fine.step(ratio)
        coarse.copy_channel_states(fine)
        coarse.step(1)
    for i in range(len(ds)):
        for f in range(2):
            np.testing.assert_array_equal(coarse.channel_states(i, f), fine.channel_states(i, f))
    feats = coarse.closure_features()
    ca = coarse.channel_ca()
    o = 0
    for i, d in enumerate(ds):
        u = coarse.species(i, 0)
        n = u.shape
        st = np.concatenate([coarse.channel_states(i, f) for f in range(2)])
        chans = np.vstack([d.geometry.boundary("ryr").pairs, d.geometry.boundary("p_ryr").pairs]).astype(int)
        for c, p in enumerate(chans):
            fr, to = p[:3], p[3:]
            want = np.empty(30)
            for k in range(27):
                q = [min(max(to[a] + (k // (1, 3, 9)[a]) % 3 - 1, 0), n[a] - 1) for a in range(3)]
                want[k] = u[q[0], q[1], q[2]]
            want[27] = u[fr[0], fr[1], fr[2]]
            cheb = np.max(np.abs(chans[:, 3:] - to), axis=1)
            want[28] = st[cheb <= 1].sum()
            want[29] = st[cheb <= 2].sum()
            np.testing.assert_allclose(feats[o + c], want, rtol=1e-6, atol=1e-30)
            assert ca[o + c] == pytest.approx(u[to[0], to[1], to[2]], rel=1e-12)
        o += len(chans)
    assert o == coarse.num_channels
    fine.close()
    coarse.close()


@pytest.fixture(scope="module")
def paired():
    train = closure.paired_samples(_dyads(range(12)), list(range(12)), coarse_steps=300, sample_every=2)
    _say(f"training set: {len(train.ca_fine)} samples from {train.fine_steps} fine / {train.coarse_steps} coarse steps, "
         f"{int(closure.flux_selection(train).sum())} with the channel open")
    return train


@pytest.fixture(scope="module")
def trained(paired):
    net, norm = closure.train(paired, mo.init_params(0), steps=2500, batch=1024, lr=2e-3)
    yield net, norm
    net.close()


@pytest.fixture(scope="module")
def trained_flux(paired, trained):
    net, norm = closure.train_flux(paired, trained[1], mo.init_params(1), steps=2500, batch=1024, lr=2e-3)
    yield net, norm
    net.close()


def test_trained_closure_beats_the_plain_coarse_grid(trained):
    """On held-out dyads, the closure's log-Ca error at the channel mouth
    is a fraction of the error of gating on the coarse cleft voxel."""
    net, norm = trained
    held = closure.paired_samples(_dyads(range(100, 106)), list(range(100, 106)), coarse_steps=200, sample_every=4)
    e_closure, e_plain = closure.log_errors(net, norm, held)
    opened = held.raw[:, 28] > 0
    _say(f"held-out log-Ca MSE: closure {e_closure:.4f}, plain coarse {e_plain:.4f} "
         f"({len(held.ca_fine)} samples, {opened.mean():.0%} with an open channel nearby)")
    assert e_closure < 0.2 * e_plain


def test_closure_gating_uses_the_network_and_monitors_its_range(trained):
    net, norm = trained
    ds = _dyads(range(200, 204))
    ens = cs.Ensemble([(d.geometry, d.model, cs.Options(h=12.0, seed=20 + k, open={"ryr": [d.open_index]}))
                       for k, d in enumerate(ds)], instances=list(range(4)))
    ens.set_gating(cs.Ensemble.GATE_CLOSURE, net)
    ens.step(50)
    # the predictions of a step come from the state and channels at its start
    feats = ens.closure_features()
    ens.step(1)
    pred = ens.closure_predictions()
    want = norm.concentration(net.forward(norm.inputs(feats)))
    assert np.all(np.isfinite(pred)) and np.all(pred > 0)
    np.testing.assert_allclose(pred, want, rtol=0.03)
    assert ens.monitor().sum() == 0, "inputs of these dyads lie inside the training range"
    # a range no input can meet: every evaluation is flagged
    tight = closure.Normalisation(norm.in_mean, norm.in_std, np.full(30, 1e9, np.float32),
                                  np.full(30, 1e9, np.float32), norm.out_mean, norm.out_std, 0.0)
    tight.apply(net)
    ens.step(3)
    flags = ens.monitor()
    per = np.bincount(ens.channel_owner(), minlength=4)
    np.testing.assert_array_equal(flags, 3 * per)
    norm.apply(net)
    ens.close()


def test_shipped_closure_on_unseen_dyads():
    """The committed parameters (train_closure.py: 48 dyads, seeds 0-47)
    on dyads of other seeds and the same size range."""
    import os
    path = os.path.join(os.path.dirname(closure.__file__), "closure_params.npz")
    net, norm, trained = closure.load(path)
    assert trained
    seeds = list(range(30000, 30006))
    held = closure.paired_samples(_dyads(seeds, edge=(256.0, 1024.0), ryr=(9, 150)), seeds,
                                  coarse_steps=200, sample_every=4)
    e_closure, e_plain = closure.log_errors(net, norm, held)
    _say(f"shipped closure, unseen dyads: log-Ca MSE closure {e_closure:.5f}, plain coarse {e_plain:.5f} "
         f"({len(held.ca_fine)} samples)")
    assert e_closure < 0.05 * e_plain
    net.close()


def test_closure_gating_replayed_as_graphs_equals_direct_launches(trained):
    """Closure gating inside captured periods (features, prediction and
    gating per gating step), replayed several times with eager work and
    synchronisations between the calls, equals direct launches."""
    net, norm = trained
    ds = _dyads(range(300, 304))
    mk = lambda: cs.Ensemble([(d.geometry, d.model, cs.Options(h=12.0, seed=30 + k, open={"ryr": [d.open_index]}))
                              for k, d in enumerate(ds)], instances=list(range(4)))
    a, b = mk(), mk()
    for e in (a, b):
        e.set_gating(cs.Ensemble.GATE_CLOSURE, net)
    a.set_graph(True)
    for rep in range(3):
        n = 40 + 3 * rep
        a.step(n)
        b.step(n)
        np.testing.assert_array_equal(a.stats()[2], b.stats()[2])
        a.synchronize()
        np.testing.assert_array_equal(a.closure_predictions(), b.closure_predictions())
        for i in range(a.count):
            for f in range(2):
                np.testing.assert_array_equal(a.channel_states(i, f), b.channel_states(i, f))
            for s in range(a.ns):
                np.testing.assert_array_equal(a.species(i, s), b.species(i, s), err_msg=f"replay {rep} member {i}")
        _say(f"closure graph replay {rep}: equal after {a.time_ind - 1} steps")
    np.testing.assert_array_equal(a.monitor(), b.monitor())
    a.close()
    b.close()


def _constant_net(value):
    """A network whose output is value for every input: zero weights, so
    every hidden unit is GELU(0) = 0 and the output is exp(out_mean)."""
    net = cs.Mlp(np.zeros(cs.MLP_NUM_PARAMS, np.float32))
    net.set_normalisation(np.zeros(30, np.float32), np.ones(30, np.float32), np.full(30, -np.inf, np.float32),
                          np.full(30, np.inf, np.float32), float(np.log(value)), 1.0, 0.0)
    return net


def test_flux_closure_scales_each_channel_release(trained):
    """One step from the same state with the same gating: release ratios
    1, 1/2 and 1/4 change the fields only at the channels' voxels, by J,
    J/2 and J/4 (J the coarse two-voxel flux); ratio 1 equals no flux
    network bit for bit."""
    net, _ = trained
    ds = _dyads(range(400, 404))
    runs = {}
    for name, ratio in (("none", None), ("r1", 1.0), ("r2", 0.5), ("r4", 0.25)):
        e = cs.Ensemble([(d.geometry, d.model, cs.Options(h=12.0, seed=40 + k, open={"ryr": [d.open_index]}))
                         for k, d in enumerate(ds)], instances=list(range(4)))
        e.set_gating(cs.Ensemble.GATE_CLOSURE, net)
        fnet = _constant_net(ratio) if ratio else None
        if fnet is not None:
            e.set_flux_closure(fnet)
        e.step(1)
        runs[name] = (e, fnet)
    e1 = runs["r1"][0]
    np.testing.assert_allclose(e1.flux_predictions(), 1.0, rtol=0, atol=0)
    moved = 0
    for i in range(len(ds)):
        for f in range(2):
            np.testing.assert_array_equal(e1.channel_states(i, f), runs["none"][0].channel_states(i, f))
        for s in range(e1.ns):
            a = runs["none"][0].species(i, s)
            np.testing.assert_array_equal(e1.species(i, s), a, err_msg=f"member {i} species {s}")
        a, b, q = (runs[k][0].species(i, 0) for k in ("r1", "r2", "r4"))
        d1, d2 = a - b, b - q
        np.testing.assert_allclose(d1, 2 * d2, rtol=1e-4, atol=1e-6 * float(np.max(np.abs(d1))))
        moved += int(np.count_nonzero(np.abs(d1) > 1e-3))
    assert moved > 0, "no channel released calcium"
    for e, fnet in runs.values():
        e.close()
        if fnet is not None:
            fnet.close()


def test_flux_predictions_are_the_network_on_the_features(trained, trained_flux):
    net, _ = trained
    fnet, fnorm = trained_flux
    ds = _dyads(range(500, 504))
    ens = cs.Ensemble([(d.geometry, d.model, cs.Options(h=12.0, seed=50 + k, open={"ryr": [d.open_index]}))
                       for k, d in enumerate(ds)], instances=list(range(4)))
    ens.set_gating(cs.Ensemble.GATE_CLOSURE, net)
    ens.set_flux_closure(fnet)
    ens.step(30)
    # every step ends with both networks evaluated on the state it leaves
    feats = ens.closure_features()
    got = ens.flux_predictions()
    want = fnorm.concentration(fnet.forward(fnorm.inputs(feats)))
    assert np.all(np.isfinite(got)) and np.all(got > 0)
    np.testing.assert_allclose(got, want, rtol=0.03)
    ens.close()


def test_trained_flux_closure_beats_the_plain_coarse_release(trained, trained_flux):
    fnet, fnorm = trained_flux
    held = closure.paired_samples(_dyads(range(100, 106)), list(range(100, 106)), coarse_steps=200, sample_every=4)
    f_c, f_p = closure.flux_errors(fnet, fnorm, held)
    sel = closure.flux_selection(held)
    _say(f"held-out MSE of the log release ratio over {int(sel.sum())} open-channel samples: flux closure "
         f"{f_c:.4f}, plain coarse {f_p:.4f} (mean ratio {np.mean(held.ratio[sel]):.3f})")
    assert f_c < 0.2 * f_p
