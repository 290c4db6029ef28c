"""Initial parameters of the closure network (30-64-64-1): He-normal
weights (standard deviation sqrt(2 / fan-in)), the output layer at
sqrt(1 / fan-in), zero biases, and zero rows for the two padding inputs.
Packed as the C ABI expects (cs_mlp_create): W1 [32][64], b1, W2 [64][64],
b2, w3, b3."""
import numpy as np

N_IN, N_PAD, N_H = 30, 32, 64


def params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, np.sqrt(2.0 / N_IN), (N_PAD, N_H))
    w1[N_IN:] = 0.0
    w2 = rng.normal(0.0, np.sqrt(2.0 / N_H), (N_H, N_H))
    w3 = rng.normal(0.0, np.sqrt(1.0 / N_H), N_H)
    return np.concatenate([w1.ravel(), np.zeros(N_H), w2.ravel(), np.zeros(N_H), w3, np.zeros(1)]).astype(np.float32)
