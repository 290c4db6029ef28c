"""Spark statistics of one calcium_sparks realisation.

The reference ships no analysis code, so these definitions are this
project's own; the same functions run on GPU realisations (csparks.
Simulation) and on CPU oracle realisations, which is what a distribution
comparison between the two needs.

  * peak dF/F0: the maximum over outputs and line positions of
    (F - F0) / F0, where F is the linescan of the indicator species
    (Species_compute_convolution_with_gaussian_function_locally,
    species.c:1960-2035) and F0 the linescan at t = 0;
  * duration: the full duration at half maximum of max_x dF/F0 over time,
    with linear interpolation at the two crossings; a signal still above
    half at the end of the run counts until the end;
  * RyRs recruited: channels of every discrete boundary that were open at
    the start or after any gating step.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence, Tuple

import numpy as np


@dataclass
class SparkStats:
    peak: float          # peak dF/F0
    duration: float      # ms
    recruited: int       # channels ever open


def fdhm(times: np.ndarray, signal: np.ndarray) -> float:
    """Full duration at half maximum of a sampled signal (0 if it never
    rises above 0)."""
    t = np.asarray(times, dtype=float)
    y = np.asarray(signal, dtype=float)
    if y.size == 0 or not np.max(y) > 0:
        return 0.0
    k = int(np.argmax(y))
    half = 0.5 * y[k]
    i = k
    while i > 0 and y[i - 1] >= half:
        i -= 1
    if i == 0:
        t0 = t[0]
    else:   # crossing between i-1 (below) and i (at or above)
        t0 = t[i - 1] + (half - y[i - 1]) / (y[i] - y[i - 1]) * (t[i] - t[i - 1])
    j = k
    while j < y.size - 1 and y[j + 1] >= half:
        j += 1
    if j == y.size - 1:
        t1 = t[-1]
    else:
        t1 = t[j] + (y[j] - half) / (y[j] - y[j + 1]) * (t[j + 1] - t[j])
    return float(t1 - t0)


def line_through(voxels: np.ndarray, domain: int) -> Tuple[int, int, int]:
    """Linescan offsets (relative to the grid centre, as -l takes them) of
    a line along x through the centre of the voxels of `domain`."""
    idx = np.argwhere(np.asarray(voxels) == domain)
    if idx.size == 0:
        raise ValueError(f"no voxel of domain {domain}")
    n = np.asarray(voxels).shape
    c = np.round(idx.mean(axis=0)).astype(int)
    return (0, int(c[1] - n[1] // 2), int(c[2] - n[2] // 2))


def discrete_count(run) -> int:
    """Number of discrete boundaries (oracle) or discrete fluxes (GPU)."""
    if hasattr(run, "setup"):
        return sum(1 for f in run.setup.problem.fluxes[:run.setup.problem.num_fluxes] if f.discrete)
    return int(run.info.num_discrete)


def realise(run, nsteps: int, save_every: int, species: int, offsets: Sequence[int],
            conv: Sequence[float], axis: int = 0, num_domains: int = 0) -> SparkStats:
    """Advance `run` (a csparks.Simulation or an oracle) nsteps steps,
    sampling the linescan every save_every steps, and return its spark
    statistics."""
    f0 = run.linescan(species, axis, offsets, conv, num_domains)
    times, peaks = [0.0], [0.0]
    done = 0
    while done < nsteps:
        k = min(save_every, nsteps - done)
        run.step(k)
        done += k
        f = run.linescan(species, axis, offsets, conv, num_domains)
        times.append(done * run.dt)
        peaks.append(float(np.max((f - f0) / f0)))
    peaks = np.asarray(peaks)
    recruited = sum(int(np.sum(run.recruited(k))) for k in range(discrete_count(run)))
    return SparkStats(peak=float(peaks.max()), duration=fdhm(np.asarray(times), peaks), recruited=recruited)


def realise_batch(runs, nsteps: int, save_every: int, species: int, offsets: Sequence[int],
                  conv: Sequence[float], axis: int = 0, num_domains: int = 0):
    """realise() for several runs at once: every run is advanced by one
    output interval before any linescan is taken, so GPU runs on separate
    contexts (streams) proceed concurrently."""
    f0 = [r.linescan(species, axis, offsets, conv, num_domains) for r in runs]
    times, peaks = [0.0], [[0.0] for _ in runs]
    done = 0
    while done < nsteps:
        k = min(save_every, nsteps - done)
        for r in runs:
            r.step(k)
        done += k
        times.append(done * runs[0].dt)
        for i, r in enumerate(runs):
            f = r.linescan(species, axis, offsets, conv, num_domains)
            peaks[i].append(float(np.max((f - f0[i]) / f0[i])))
    out = []
    for i, r in enumerate(runs):
        p = np.asarray(peaks[i])
        rec = sum(int(np.sum(r.recruited(k))) for k in range(discrete_count(r)))
        out.append(SparkStats(peak=float(p.max()), duration=fdhm(np.asarray(times), p), recruited=rec))
    return out


def realise_ensemble(ens, nsteps: int, save_every: int, species: int, offsets: Sequence[int],
                     conv: Sequence[float], axis: int = 0, num_domains: int = 0, progress=None):
    """realise() for every member of a csparks.Ensemble: the members step
    together and one batched linescan per output serves them all.
    progress(done, nsteps), if given, is called after every output."""
    f0 = ens.linescans(species, axis, offsets, conv, num_domains)
    times, peaks = [0.0], [[0.0] for _ in range(ens.count)]
    done = 0
    while done < nsteps:
        k = min(save_every, nsteps - done)
        ens.step(k)
        done += k
        times.append(done * ens.dt)
        for i, f in enumerate(ens.linescans(species, axis, offsets, conv, num_domains)):
            peaks[i].append(float(np.max((f - f0[i]) / f0[i])))
        if progress is not None:
            progress(done, nsteps)
    _, _, recruited = ens.summary()
    return [SparkStats(peak=float(np.max(p)), duration=fdhm(np.asarray(times), np.asarray(p)),
                       recruited=int(r)) for p, r in zip(peaks, recruited)]
