"""GPU parity of the closure network (cs_mlp_*) against its CPU FP64
oracle (oracle/mlp_oracle.py, pinned by finite differences): the oracle
rounds to bfloat16 where the device holds bfloat16 operands, so the
device's FP32 accumulation is the remaining difference -- outputs agree to
1e-3 of their scale, gradients to 2e-3 in norm (a bfloat16 rounding that
FP32 and FP64 place on opposite sides of a tie moves one element by 2^-8).
Against the unrounded FP64 network the bound is the bfloat16 one, 5e-2.
Training must reduce a regression loss, and one Adam step must equal the
oracle's update."""
import sys

import numpy as np
import pytest

import csparks as cs
import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(seed, n):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed) + rng.normal(0, 0.05, mo.N_PARAMS)
    p[mo.N_IN * mo.N_H:mo.N_PAD * mo.N_H] = 0.0     # padding rows of W1
    x = rng.normal(0, 1.0, (n, mo.N_IN))
    t = rng.normal(0, 1.0, n)
    return p.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64), \
        t.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [1, 31, 32, 1000, 4099])
def test_forward_matches_oracle(n):
    p, x, _ = _case(1, n)
    net = cs.Mlp(p)
    y = net.forward(x)
    yb, _ = mo.forward(p, x, bf16_ops=True)
    y64, _ = mo.forward(p, x)
    scale = max(1.0, float(np.max(np.abs(yb))))
    err, err64 = float(np.max(np.abs(y - yb))), float(np.max(np.abs(y - y64)))
    _say(f"n={n}: |y - oracle(bf16)| {err:.2e}, |y - oracle(fp64)| {err64:.2e}, scale {scale:.2f}")
    assert err <= 1e-3 * scale
    assert err64 <= 5e-2 * scale
    net.close()


@pytest.mark.parametrize("n", [64, 1000, 5000])
def test_gradient_matches_oracle(n):
    p, x, t = _case(2, n)
    net = cs.Mlp(p)
    g, loss = net.grad(x, t)
    gb, lb = mo.grad(p, x, t, bf16_ops=True)
    g64, _ = mo.grad(p, x, t)
    rel = float(np.linalg.norm(g - gb) / np.linalg.norm(gb))
    rel64 = float(np.linalg.norm(g - g64) / np.linalg.norm(g64))
    _say(f"n={n}: grad rel err vs oracle(bf16) {rel:.2e}, vs oracle(fp64) {rel64:.2e}; loss {loss:.6f} vs {lb:.6f}")
    assert rel <= 2e-3
    assert rel64 <= 5e-2
    assert abs(loss - lb) <= 1e-3 * lb
    # every parameter block carries its share (no block left at zero)
    for name, blk in mo.unpack(g).items():
        ref = mo.unpack(gb)[name]
        if name == "W1":
            blk, ref = blk[:mo.N_IN], ref[:mo.N_IN]
        assert np.linalg.norm(blk - ref) <= 5e-3 * np.linalg.norm(ref) + 1e-7, name
    net.close()


def test_one_adam_step_equals_oracle_update():
    p, x, t = _case(3, 2048)
    net = cs.Mlp(p)
    net.set_data(x, t, seed=1)
    net.train(1, 2048, 1e-3)                 # one full batch: the order does not matter
    gb, _ = mo.grad(p, x, t, bf16_ops=True)
    pr, _, _ = mo.adam(p, gb, np.zeros_like(p), np.zeros_like(p), 1, 1e-3)
    q = net.params.astype(np.float64)
    big = np.abs(gb) > 1e-2 * np.max(np.abs(gb))
    np.testing.assert_allclose(q[big], pr[big], rtol=0, atol=2e-6)
    # Adam's first step moves a weight by at most lr, up to the FP32 rounding of
    # the bias corrections 1 - 0.9 and 1 - 0.999 (a few 1e-5 relative)
    assert np.max(np.abs(q - p)) <= 1e-3 * (1 + 1e-4)
    net.close()


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(4)
    x = rng.uniform(-2, 2, (40000, mo.N_IN))
    t = np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2 + 0.1 * x[:, 4:].sum(axis=1)
    net = cs.Mlp(mo.init_params(4))
    _, l0 = net.grad(x[:4096], t[:4096])
    net.set_data(x, t, seed=2)
    for epoch in range(6):
        l = net.train(150, 1024, 2e-3)
        _say(f"epoch {epoch}: batch loss {l:.5f}")
    _, l1 = net.grad(x[:4096], t[:4096])
    _say(f"loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.05 * l0
    net.close()
