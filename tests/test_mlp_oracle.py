"""Pins of the closure network's CPU oracle (oracle/mlp_oracle.py): its
gradient against central finite differences of its own loss in FP64, GELU
against the normal CDF, the bfloat16 rounding against its definition, and
Adam's first step against its closed form."""
import numpy as np
import pytest
from scipy.stats import norm

import mlp_oracle as mo


def _case(seed=0, n=48):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p += rng.normal(0, 0.05, p.shape)          # non-zero biases too
    x = rng.normal(0, 1.0, (n, mo.N_IN))
    t = rng.normal(0, 1.0, n)
    return p, x, t


def test_gradient_matches_central_differences():
    p, x, t = _case()
    g, _ = mo.grad(p, x, t)
    rng = np.random.default_rng(1)
    # every bias and output weight, and a sample of each weight matrix
    idx = list(rng.choice(mo.N_PAD * mo.N_H, 40, replace=False))
    o = mo.N_PAD * mo.N_H
    idx += list(range(o, o + mo.N_H))
    idx += list(o + mo.N_H + rng.choice(mo.N_H * mo.N_H, 40, replace=False))
    idx += list(range(mo.N_PARAMS - 2 * mo.N_H - 1, mo.N_PARAMS))
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (mo.loss(p + e, x, t) - mo.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_padding_inputs_have_zero_gradient():
    p, x, t = _case(3)
    g, _ = mo.grad(p, x, t)
    W1 = mo.unpack(g)["W1"]
    assert np.all(W1[mo.N_IN:, :] == 0.0)


def test_gelu_is_z_times_normal_cdf():
    z = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(mo.gelu(z), z * norm.cdf(z), rtol=1e-12, atol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(mo.gelu_grad(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even (1.0)
    assert mo.bf16(1.0 + 2.0 ** -8)[()] == 1.0
    assert mo.bf16(1.0 + 3 * 2.0 ** -8)[()] == 1.0 + 2.0 ** -6
    assert mo.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12)[()] == 1.0 + 2.0 ** -7
    x = np.random.default_rng(0).normal(0, 100, 1000)
    r = mo.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    np.testing.assert_array_equal(mo.bf16(r), r)       # idempotent


def test_adam_first_step_moves_each_weight_by_lr():
    p, x, t = _case(5)
    g, _ = mo.grad(p, x, t)
    p1, m, v = mo.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1, 1e-3)
    moved = p - p1
    nz = np.abs(g) > 1e-3          # |step - lr| <= lr eps / |g|
    assert nz.sum() > 100
    np.testing.assert_allclose(moved[nz], 1e-3 * np.sign(g[nz]), rtol=1e-4)


@pytest.mark.parametrize("seed", [0, 7])
def test_bf16_forward_stays_near_fp64(seed):
    p, x, t = _case(seed)
    y64, _ = mo.forward(p, x)
    y16, _ = mo.forward(p, x, bf16_ops=True)
    assert np.max(np.abs(y16 - y64)) < 0.05 * max(1.0, np.max(np.abs(y64)))


def test_device_erf_constants_meet_their_published_bound():
    """The device GELU takes erf from Abramowitz and Stegun 7.1.26
    (|error| <= 1.5e-7).  The constants are read from the device source
    (cs_mlp.h, mlp_erf) and evaluated in the same order in FP64, so a
    mistyped digit or a term out of order fails against scipy's erf."""
    import os
    import re
    from scipy.special import erf
    from conftest import REPO
    src = open(os.path.join(REPO, "hpc-ca-phospattern_amd", "csrc", "cs_mlp.h")).read()
    body = src[src.index("__device__ inline float mlp_erf"):]
    body = body[:body.index("\n}\n")]
    num = r"(-?\d+\.\d+)f"
    p = float(re.search(r"fmaf\(" + num + r", ax, 1\.0f\)", body).group(1))
    horner = [float(re.search(r"float p = " + num, body).group(1))]
    horner += [float(v) for v in re.findall(r"p = fmaf\(p, t, " + num + r"\)", body)]
    assert len(horner) == 5
    x = np.linspace(-6.0, 6.0, 200001)
    ax = np.abs(x)
    t = 1.0 / (1.0 + p * ax)
    poly = horner[0]
    for a in horner[1:]:
        poly = poly * t + a
    approx = np.copysign(1.0 - poly * t * np.exp(-ax * ax), x)
    err = float(np.max(np.abs(approx - erf(x))))
    assert err <= 1.5e-7, err
