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

Training data come from paired runs of the same 12 nm dyads at 3 nm (fine)
and 12 nm (coarse): the fine ensemble gates with Ca-independent rates that
visit many open/closed configurations, and the coarse ensemble replays its
channel states (GATE_REPLAY), so both grids see the same release schedule;
each coarse step's features are paired with the fine mouth concentration
at the same time.  Inputs: the coarse Ca of the 3x3x3 voxels around the
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


def paired_samples(dyads: Sequence, seeds: Sequence[int], fine_h: float = 3.0, coarse_h: float = 12.0,
                   coarse_steps: int = 400, sample_every: int = 2, warmup: int = 10,
                   kinetics: Tuple[float, float] = (20.0, 40.0), precision: str = "f32") -> PairedSamples:
    """Run the dyads (geometries at coarse_h) at fine_h and coarse_h side by
    side with one channel schedule and collect (features, labels)."""
    members = []
    for d, s in zip(dyads, seeds):
        m = schedule_model(d.model, *kinetics)
        members.append((d.geometry, m, s, d.open_index))
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
    raw, yf, yc, own = [], [], [], []
    for k in range(1, coarse_steps + 1):
        fine.step(steps_per)
        coarse.copy_channel_states(fine)
        coarse.step(1)
        if k >= warmup and (k - warmup) % sample_every == 0:
            raw.append(coarse.closure_features())
            yf.append(fine.channel_ca())
            yc.append(coarse.channel_ca())
            own.append(owner)
    fine.close()
    coarse.close()
    return PairedSamples(np.concatenate(raw), np.concatenate(yf), np.concatenate(yc), np.concatenate(own),
                         coarse_steps * steps_per, coarse_steps)


def train(samples: PairedSamples, params0: np.ndarray, steps: int = 3000, batch: int = 1024, lr: float = 2e-3,
          seed: int = 0, margin: float = 0.5):
    """Fit the network to the paired samples on the GPU; returns the
    network (with its normalisation applied) and the normalisation."""
    norm = fit_normalisation(samples.raw, samples.ca_fine, margin)
    net = cs.Mlp(params0)
    norm.apply(net)
    net.set_data(norm.inputs(samples.raw), norm.target(samples.ca_fine), seed=seed)
    left = steps
    while left > 0:
        k = min(left, 500)
        net.train(k, min(batch, len(samples.ca_fine)), lr)
        left -= k
    return net, norm


def log_errors(net: cs.Mlp, norm: Normalisation, samples: PairedSamples) -> Tuple[float, float]:
    """Mean squared error of log Ca at the channel mouth: the closure's
    prediction, and the coarse cleft voxel's Ca (no closure)."""
    pred = norm.concentration(net.forward(norm.inputs(samples.raw)))
    t = np.log(np.maximum(samples.ca_fine, LOG_FLOOR))
    e_closure = float(np.mean((np.log(pred) - t) ** 2))
    e_plain = float(np.mean((np.log(np.maximum(samples.ca_coarse, LOG_FLOOR)) - t) ** 2))
    return e_closure, e_plain


def save(path: str, net: cs.Mlp, norm: Normalisation, **meta):
    """Trained parameters and normalisation as a plain .npz (no pickles)."""
    extra = {k: np.asarray(v) for k, v in meta.items()}
    np.savez(path, params=net.params, in_mean=norm.in_mean, in_std=norm.in_std, in_lo=norm.in_lo, in_hi=norm.in_hi,
             out=np.array([norm.out_mean, norm.out_std, norm.margin]), **extra)


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
