"""The learned closure of the coarse-grid method: paired fine/coarse runs,
feature normalisation, training.

The reference gates each channel on the calcium concentration of its cleft
voxel (evaluate_RyR_stochastically, boundaryfluxes.c:295-359; the voxel is
the channel's `to` voxel, species.c:2569).  On a 12 nm grid that voxel
averages the nanodomain a channel sees at 3 nm over 64 fine voxels, so the
coarse grid under-recruits.  The closure predicts, from coarse-grid
features, the concentration the fine grid puts at the channel mouth -- the
channel's fine `to` voxel, which under the reference's refinement of one
12 nm geometry (species.c:1003-1020: a discrete boundary maps to the
centre sub-voxel pair) is the centre sub-voxel of the coarse one -- and the
coarse ensemble gates on that prediction (Ensemble.set_gating).

A second network on the same inputs closes the release: the fine grid's
two-voxel flux (ryr_flux, boundaryfluxes.c:87-141) acts between 3 nm
voxels at the mouth, which deplete on the SR side and fill on the cleft
side within a few fine steps, while the coarse grid's acts between 12 nm
voxels that barely move; the coarse grid then releases several times too
much calcium per opening.  The flux network predicts, for an open channel,
the ratio of the fine release to the coarse one over the same time
(Ensemble.set_flux_closure), and the coarse flux is scaled by it.

Training data come from paired runs of the same 12 nm dyads at 3 nm (fine)
and 12 nm (coarse): the fine ensemble gates with Ca-independent rates that
visit many open/closed configurations, and the coarse ensemble replays its
channel states (GATE_REPLAY), so both grids see the same release schedule;
each coarse step's features are paired with the fine mouth concentration
at the same time, and for the open channels with the release ratio.  Inputs: the coarse Ca of the 3x3x3 voxels around the
cleft voxel, the SR Ca at the channel's SR voxel, and the open channels
within 1 and 2 coarse voxels (Ensemble.closure_features); concentrations
enter as logarithms, everything standardised on the training set, and the
target is the standardised logarithm of the fine mouth Ca.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

import csparks as cs

N_CONC = 28            # concentration inputs (27 cleft neighbourhood + SR)
LOG_FLOOR = 1e-6       # uM, floor under the logarithm (T-tubule voxels hold no calcium)


def schedule_model(model: cs.SpeciesModel, k_open: float = 20.0, k_close: float = 40.0) -> cs.SpeciesModel:
    """The same model with Ca-independent gating (constant rates, per ms):
    the clamped power law of the reference with exponent 0 is the constant
    max(min(1, k_max), k_min), so k_min = k_max = k gives rate k."""
    def const(ch: Optional[cs.ChannelModel]):
        if ch is None:
            return None
        return dataclasses.replace(ch, Kd_open=1.0, n_open=0.0, k_min_open=k_open, k_max_open=k_open,
                                   Kd_close=1.0, n_close=0.0, k_min_close=k_close, k_max_close=k_close)
    return dataclasses.replace(model, ryr=const(model.ryr), p_ryr=const(model.p_ryr))


def transform(raw: np.ndarray) -> np.ndarray:
    """Raw features -> the network's unstandardised inputs (logarithm of the
    concentrations, counts as they are), in float32 like the device."""
    f = np.asarray(raw, dtype=np.float32).copy()
    f[:, :N_CONC] = np.log(np.maximum(f[:, :N_CONC], np.float32(LOG_FLOOR)))
    return f


@dataclass
class Normalisation:
    in_mean: np.ndarray
    in_std: np.ndarray
    in_lo: np.ndarray
    in_hi: np.ndarray
    out_mean: float
    out_std: float
    margin: float = 0.5

    def inputs(self, raw: np.ndarray) -> np.ndarray:
        return ((transform(raw) - self.in_mean) / self.in_std).astype(np.float32)

    def target(self, ca: np.ndarray) -> np.ndarray:
        return ((np.log(np.maximum(ca, LOG_FLOOR)) - self.out_mean) / self.out_std).astype(np.float32)

    def concentration(self, y: np.ndarray) -> np.ndarray:
        return np.exp(np.asarray(y, dtype=np.float64) * self.out_std + self.out_mean)

    def apply(self, net: cs.Mlp):
        net.set_normalisation(self.in_mean, self.in_std, self.in_lo, self.in_hi, self.out_mean, self.out_std,
                              self.margin)


def fit_normalisation(raw: np.ndarray, ca_fine: np.ndarray, margin: float = 0.5) -> Normalisation:
    f = transform(raw).astype(np.float64)
    mean = f.mean(axis=0)
    std = np.maximum(f.std(axis=0), 1e-3)
    x = (f - mean) / std
    t = np.log(np.maximum(ca_fine, LOG_FLOOR))
    return Normalisation(mean.astype(np.float32), std.astype(np.float32), x.min(axis=0).astype(np.float32),
                         x.max(axis=0).astype(np.float32), float(t.mean()), float(max(t.std(), 1e-3)), margin)


@dataclass
class PairedSamples:
    raw: np.ndarray        # [n, 30] coarse features
    ca_fine: np.ndarray    # [n] fine Ca at the channel mouth
    ca_coarse: np.ndarray  # [n] coarse Ca at the cleft voxel (what plain coarse gating reads)
    owner: np.ndarray      # [n] realisation of the sample
    fine_steps: int
    coarse_steps: int
    open: Optional[np.ndarray] = None    # [n] channel open for the next step's flux
    ratio: Optional[np.ndarray] = None   # [n] fine release / coarse release (where open)


def release_factors(ens: cs.Ensemble, h: float) -> np.ndarray:
    """Per channel (channel order), the two-voxel flux per step and unit
    concentration difference of its flux: (1 - K_0) / 2 of ryr_flux
    (boundaryfluxes.c:87-141) at the ensemble's dt and voxel size h."""
    lib = cs.device_lib()
    out = []
    for st in ens.setups:
        p = st.problem
        for f in range(p.num_fluxes):
            F = p.fluxes[f]
            if F.discrete:
                k = float(lib.cs_ryr_flux(ens.dt, h, 1.0, 0.0, F.params))
                out.extend([k] * int(F.num_pairs))
    return np.asarray(out)


def release_ratio(fine_from, fine_to, coarse_from, coarse_to, k_fine, k_coarse, steps_per: int,
                  h_ratio: float) -> np.ndarray:
    """Fine release over one coarse step (steps_per fine steps of
    k_fine (from - to) into a voxel of (1/h_ratio)^3 the volume) over the
    coarse grid's k_coarse (from - to) for the same channel."""
    d_f = np.asarray(fine_from) - np.asarray(fine_to)
    d_c = np.maximum(np.asarray(coarse_from) - np.asarray(coarse_to), 1e-3)
    return np.maximum(d_f, 0.0) * k_fine * steps_per / (d_c * k_coarse * h_ratio ** 3)


def paired_samples(dyads: Sequence, seeds: Sequence[int], fine_h: float = 3.0, coarse_h: float = 12.0,
                   coarse_steps: int = 400, sample_every: int = 2, warmup: int = 10,
                   kinetics: Tuple[float, float] = (20.0, 40.0), precision: str = "f32",
                   say=None) -> PairedSamples:
    """Run the dyads (geometries at coarse_h) at fine_h and coarse_h side by
    side with one channel schedule and collect (features, labels); say,
    if given, receives a progress line every 50 coarse steps."""
    members = []
    for d, s in zip(dyads, seeds):
        m = schedule_model(d.model, *kinetics)
        members.append((d.geometry, m, s, d.open_index))
    if say is not None:
        say(f"building the {fine_h:g} nm and {coarse_h:g} nm ensembles of {len(members)} dyads")
    fine = cs.Ensemble([(g, m, cs.Options(h=fine_h, seed=s, open={"ryr": [k]}, precision=precision))
                        for g, m, s, k in members], instances=list(range(len(members))))
    coarse = cs.Ensemble([(g, m, cs.Options(h=coarse_h, seed=s, open={"ryr": [k]}, precision=precision))
                          for g, m, s, k in members], instances=list(range(len(members))))
    ratio = coarse.dt / fine.dt
    steps_per = int(round(ratio))
    if abs(ratio - steps_per) > 1e-9 * ratio:
        raise ValueError(f"coarse dt is not a whole number of fine steps ({ratio})")
    coarse.set_gating(cs.Ensemble.GATE_REPLAY)
    coarse.copy_channel_states(fine)
    owner = coarse.channel_owner()
    kf, kc = release_factors(fine, fine_h), release_factors(coarse, coarse_h)
    raw, yf, yc, own, op, rr = [], [], [], [], [], []
    if say is not None:
        _, nvf, _ = fine.info()
        say(f"paired ensembles: {len(members)} dyads, {nvf} fine voxels, {steps_per} fine steps per coarse step")
    for k in range(1, coarse_steps + 1):
        fine.step(steps_per)
        coarse.copy_channel_states(fine)
        coarse.step(1)
        if say is not None and k % 50 == 0:
            say(f"  coarse step {k}/{coarse_steps}")
        if k >= warmup and (k - warmup) % sample_every == 0:
            raw.append(coarse.closure_features())
            yf.append(fine.channel_ca())
            yc.append(coarse.channel_ca())
            own.append(owner)
            op.append(fine.channel_open())
            rr.append(release_ratio(fine.channel_sr_ca(), yf[-1], coarse.channel_sr_ca(), yc[-1], kf, kc,
                                    steps_per, coarse_h / fine_h))
    fine.close()
    coarse.close()
    return PairedSamples(np.concatenate(raw), np.concatenate(yf), np.concatenate(yc), np.concatenate(own),
                         coarse_steps * steps_per, coarse_steps, np.concatenate(op), np.concatenate(rr))


def train(samples: PairedSamples, params0: np.ndarray, steps: int = 3000, batch: int = 1024, lr: float = 2e-3,
          seed: int = 0, margin: float = 0.5):
    """Fit the network to the paired samples on the GPU; returns the
    network (with its normalisation applied) and the normalisation."""
    norm = fit_normalisation(samples.raw, samples.ca_fine, margin)
    net = cs.Mlp(params0)
    norm.apply(net)
    _fit(net, norm.inputs(samples.raw), norm.target(samples.ca_fine), steps, batch, lr, seed)
    return net, norm


def _fit(net: cs.Mlp, x: np.ndarray, t: np.ndarray, steps: int, batch: int, lr: float, seed: int):
    net.set_data(x, t, seed=seed)
    left = steps
    while left > 0:
        k = min(left, 500)
        net.train(k, min(batch, len(t)), lr)
        left -= k


def flux_selection(samples: PairedSamples) -> np.ndarray:
    """Samples that train the flux network: open channels with a release."""
    return (samples.open > 0) & np.isfinite(samples.ratio) & (samples.ratio > 0)


def train_flux(samples: PairedSamples, norm: Normalisation, params0: np.ndarray, steps: int = 3000,
               batch: int = 1024, lr: float = 2e-3, seed: int = 1):
    """Fit the flux network to the open channels' release ratios, on the
    gating network's input standardisation (one range monitor serves
    both); returns the network and its normalisation."""
    sel = flux_selection(samples)
    if not sel.any():
        raise ValueError("no open channel among the samples")
    t = np.log(samples.ratio[sel])
    nf = Normalisation(norm.in_mean, norm.in_std, norm.in_lo, norm.in_hi, float(t.mean()),
                       float(max(t.std(), 1e-3)), norm.margin)
    net = cs.Mlp(params0)
    nf.apply(net)
    _fit(net, nf.inputs(samples.raw[sel]), nf.target(samples.ratio[sel]), steps, batch, lr, seed)
    return net, nf


def flux_errors(net: cs.Mlp, norm: Normalisation, samples: PairedSamples) -> Tuple[float, float]:
    """Mean squared error of the log release ratio over the open channels:
    the flux network's, and the plain coarse grid's (ratio 1)."""
    sel = flux_selection(samples)
    t = np.log(samples.ratio[sel])
    pred = np.log(norm.concentration(net.forward(norm.inputs(samples.raw[sel]))))
    return float(np.mean((pred - t) ** 2)), float(np.mean(t ** 2))


def log_errors(net: cs.Mlp, norm: Normalisation, samples: PairedSamples) -> Tuple[float, float]:
    """Mean squared error of log Ca at the channel mouth: the closure's
    prediction, and the coarse cleft voxel's Ca (no closure)."""
    pred = norm.concentration(net.forward(norm.inputs(samples.raw)))
    t = np.log(np.maximum(samples.ca_fine, LOG_FLOOR))
    e_closure = float(np.mean((np.log(pred) - t) ** 2))
    e_plain = float(np.mean((np.log(np.maximum(samples.ca_coarse, LOG_FLOOR)) - t) ** 2))
    return e_closure, e_plain


def save(path: str, net: cs.Mlp, norm: Normalisation, flux=None, **meta):
    """Trained parameters and normalisation as a plain .npz (no pickles);
    flux = (network, normalisation) of the flux closure, if trained."""
    extra = {k: np.asarray(v) for k, v in meta.items()}
    if flux is not None:
        fnet, fnorm = flux
        extra["flux_params"] = fnet.params
        extra["flux_out"] = np.array([fnorm.out_mean, fnorm.out_std, fnorm.margin])
    np.savez(path, params=net.params, in_mean=norm.in_mean, in_std=norm.in_std, in_lo=norm.in_lo, in_hi=norm.in_hi,
             out=np.array([norm.out_mean, norm.out_std, norm.margin]), **extra)


def load_flux(path: str):
    """(network, normalisation) of the flux closure saved with the gating
    network, or None when the file has none."""
    import os
    if not os.path.exists(path):
        return None
    z = np.load(path)
    if "flux_params" not in z.files:
        return None
    o = z["flux_out"]
    norm = Normalisation(z["in_mean"], z["in_std"], z["in_lo"], z["in_hi"], float(o[0]), float(o[1]), float(o[2]))
    net = cs.Mlp(z["flux_params"])
    norm.apply(net)
    return net, norm


def load(path: str, init_seed: int = 0):
    """(network, normalisation, trained) from a file written by save(); an
    untrained network with a neutral normalisation when there is none (same
    cost per evaluation, for timing)."""
    import os
    if os.path.exists(path):
        z = np.load(path)
        norm = Normalisation(z["in_mean"], z["in_std"], z["in_lo"], z["in_hi"], float(z["out"][0]),
                             float(z["out"][1]), float(z["out"][2]))
        net = cs.Mlp(z["params"])
        norm.apply(net)
        return net, norm, True
    import mlp_init
    norm = Normalisation(np.zeros(cs.MLP_IN, np.float32), np.ones(cs.MLP_IN, np.float32),
                         np.full(cs.MLP_IN, -np.inf, np.float32), np.full(cs.MLP_IN, np.inf, np.float32), 0.0, 1.0, 0.0)
    net = cs.Mlp(mlp_init.params(init_seed))
    norm.apply(net)
    return net, norm, False
