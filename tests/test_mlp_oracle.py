"""Pins of the CPU MLP oracle: its analytic gradients against central finite differences (float64),
and the bfloat16 rounding against the IEEE definition (round to nearest, ties to even)."""
import numpy as np

from oracle import mlp


def small_params(rng, k=8, h=6, n=5):
    return [rng.normal(0, 0.5, (h, k)), rng.normal(0, 0.1, h), rng.normal(0, 0.5, (h, h)), rng.normal(0, 0.1, h),
            rng.normal(0, 0.5, (n, h)), rng.normal(0, 0.1, n)]


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    P = small_params(rng)
    x = rng.normal(size=(7, 8))
    t = rng.normal(size=(7, 5))
    _, grads = mlp.loss_and_grads(P, x, t)
    eps = 1e-6
    for pi, (p, g) in enumerate(zip(P, grads)):
        flat = p.reshape(-1)
        for j in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[j]
            flat[j] = old + eps
            lp, _ = mlp.loss_and_grads(P, x, t)
            flat[j] = old - eps
            lm, _ = mlp.loss_and_grads(P, x, t)
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g.reshape(-1)[j]) <= 1e-7 + 1e-5 * abs(fd), (pi, j, fd, g.reshape(-1)[j])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40], dtype=np.float64)
    r = mlp.to_bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0          # 2^-8 is the half-way point of the 7-bit mantissa
    assert r[1] == 1.0                                    # tie -> even (mantissa 0)
    assert r[2] == 1.0                                    # below half
    assert r[3] == 1.0 + 2 ** -7                          # above half
    assert r[4] == -2.5
    # every result is representable in bfloat16: low 16 bits of the float32 pattern are zero
    assert np.all((r.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)
