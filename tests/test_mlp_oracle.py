"""The CPU oracle of the multilayer perceptron (oracle/mlp_oracle.py),
pinned by mathematics: its analytic gradients against central finite
differences of its own float64 loss, GELU and its derivative against
closed forms and numerical differentiation, bfloat16 rounding against
hand-worked values, and Adam's first step against its closed form.
No GPU needed."""


import numpy as np

from oracle import mlp_oracle as mo


def test_bf16_rounding_hand_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 1.0 + 2.0 ** -7 + 2.0 ** -8, 3.0e38],
                 dtype=np.float32)
    # 3e38 is 0x7F61B1E6 in float32: rounds up to 0x7F62 = 2^127 * (1 + 98/128)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 1.0 + 2.0 ** -6, 2.0 ** 127 * (1 + 98 / 128)])
    got = mo.bf16(x)
    assert np.all(np.abs(got - want) <= 1e-7 * np.abs(want))


def test_gelu_closed_forms_and_derivative():
    assert mo.gelu(np.array(0.0)) == 0.0
    assert abs(mo.gelu(np.array(10.0)) - 10.0) < 1e-12 and abs(mo.gelu(np.array(-10.0))) < 1e-12
    assert abs(mo.gelu_grad(np.array(0.0)) - 0.5) < 1e-15
    # GELU(1) = Phi(1) = 0.8413447460685429 (standard normal CDF)
    assert abs(mo.gelu(np.array(1.0)) - 0.8413447460685429) < 1e-15
    x = np.linspace(-4, 4, 41)
    h = 1e-5
    num = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(num - mo.gelu_grad(x))) < 1e-9


def test_gradients_match_finite_differences():
    widths = [3, 8, 7, 5]
    rng = np.random.default_rng(1)
    params = [(rng.standard_normal((o, i)), rng.standard_normal(o) * 0.1) for i, o in zip(widths[:-1], widths[1:])]
    x = rng.standard_normal((4, 3))
    t = rng.standard_normal((4, 5))
    loss, grads = mo.loss_and_grads(params, x, t, bf16_ops=False)
    h = 1e-6
    for l, (w, b) in enumerate(params):
        for arr, g in ((w, grads[l][0]), (b, grads[l][1])):
            num = np.zeros_like(arr)
            for idx in np.ndindex(arr.shape):
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = mo.loss_and_grads(params, x, t, bf16_ops=False)
                arr[idx] = old - h
                lm, _ = mo.loss_and_grads(params, x, t, bf16_ops=False)
                arr[idx] = old
                num[idx] = (lp - lm) / (2 * h)
            assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_gradient_check_catches_a_dropped_term():
    """The finite-difference check above detects a plausible slip: the
    backward pass without the GELU derivative fails it."""
    widths = [3, 6, 4]
    rng = np.random.default_rng(2)
    params = [(rng.standard_normal((o, i)), np.zeros(o)) for i, o in zip(widths[:-1], widths[1:])]
    x, t = rng.standard_normal((4, 3)), rng.standard_normal((4, 4))
    _, good = mo.loss_and_grads(params, x, t, bf16_ops=False)
    saved = mo.gelu_grad
    try:
        mo.gelu_grad = lambda z: np.ones_like(z)
        _, bad = mo.loss_and_grads(params, x, t, bf16_ops=False)
    finally:
        mo.gelu_grad = saved
    assert np.max(np.abs(bad[0][0] - good[0][0])) > 1e-3


def test_adam_first_step_closed_form():
    """With m0 = v0 = 0 the bias-corrected first step is lr * g / (|g| + eps)."""
    params = [(np.array([[1.0, -2.0]]), np.array([0.5]))]
    grads = [(np.array([[0.3, -4.0]]), np.array([1e-3]))]
    new = mo.adam_step(params, grads, {}, lr=0.01)
    want_w = np.array([[1.0 - 0.01 * 0.3 / (0.3 + 1e-8), -2.0 + 0.01 * 4.0 / (4.0 + 1e-8)]])
    assert np.allclose(new[0][0], want_w, rtol=0, atol=1e-12)
    assert abs(new[0][1][0] - (0.5 - 0.01 * 1e-3 / (1e-3 + 1e-8))) < 1e-12

