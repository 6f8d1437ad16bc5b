"""Pins of the radial-surrogate network restatement (CPU)."""
import numpy as np

from oracle import mlp_oracle as mo


def _setup(H=8, O=5, B=13, seed=3):
    rng = np.random.default_rng(seed)
    p = mo.init_params(H, O, rng)
    x = rng.uniform(-1, 1, B)
    t = rng.normal(size=(B, O))
    dt = rng.normal(size=(B, O))
    wv = rng.uniform(0.5, 1.5, O)
    wd = rng.uniform(0.1, 2.0, O)
    return rng, p, x, t, dt, wv, wd, H, O


def test_tangent_matches_finite_difference_in_x():
    _, p, x, *_rest = _setup()
    H, O = 8, 5
    h = 1e-6
    y_p, _, _ = mo.forward(p, x + h, H, O)
    y_m, _, _ = mo.forward(p, x - h, H, O)
    _, dy, _ = mo.forward(p, x, H, O)
    np.testing.assert_allclose(dy, (y_p - y_m) / (2 * h), rtol=1e-7, atol=1e-8)


def test_gradient_matches_central_differences():
    _, p, x, t, dt, wv, wd, H, O = _setup()
    L, g = mo.loss_and_grad(p, x, t, dt, wv, wd, H, O)
    fd = np.empty_like(p)
    h = 1e-6
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mo.loss_and_grad(p + e, x, t, dt, wv, wd, H, O)[0] -
                 mo.loss_and_grad(p - e, x, t, dt, wv, wd, H, O)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    assert abs(L - (1 / (O * x.size)) * ((wv * (mo.forward(p, x, H, O)[0] - t) ** 2).sum()
                                         + (wd * (mo.forward(p, x, H, O)[1] - dt) ** 2).sum())) < 1e-12


def test_planted_sign_error_in_gradient_is_detected():
    """A wrong sign in the tangent-path term must break the finite-difference check."""
    _, p, x, t, dt, wv, wd, H, O = _setup()
    _, g = mo.loss_and_grad(p, x, t, dt, wv, wd, H, O)
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, H, O)
    # recompute gW1 with the -2 h1 W1 gd1 term's sign flipped
    y, dy, (h1, d1, h2, d2, t2) = mo.forward(p, x, H, O)
    sc = 1.0 / (O * x.size)
    ey, ed = 2 * sc * wv * (y - t), 2 * sc * wd * (dy - dt)
    gh2, gd2 = ey @ W3, ed @ W3
    u2 = 1 - h2 * h2
    gt2 = u2 * gd2
    ga2 = u2 * (gh2 - 2 * h2 * t2 * gd2)
    gh1, gd1 = ga2 @ W2, gt2 @ W2
    u1 = 1 - h1 * h1
    wrong = (u1 * (gh1 + 2 * h1 * W1 * gd1) * x[:, None]).sum(0) + (u1 * gd1).sum(0)
    assert np.abs(wrong - g[:H]).max() > 1e-6


def test_adam_first_step_moves_by_lr_times_sign():
    """With bias correction the first Adam step is lr * g / (|g| + eps) per coordinate."""
    rng = np.random.default_rng(0)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    q, m, v = mo.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    np.testing.assert_allclose(p - q, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    # second step against the closed form
    g2 = rng.normal(size=50)
    q2, m2, v2 = mo.adam_step(q, g2, m, v, 2, 1e-3)
    mh = (0.9 * 0.1 * g + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g * g + 0.001 * g2 * g2) / (1 - 0.999 ** 2)
    np.testing.assert_allclose(q - q2, 1e-3 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-10)


def test_refit_output_recovers_a_representable_target():
    """Targets produced by the network itself are fitted exactly by the output-layer refit."""
    rng, p, x, _, _, wv, wd, H, O = _setup(B=200)
    target_p = p.copy()
    _, _, _, _, W3, b3 = mo.unpack(target_p, H, O)
    W3[:] = rng.normal(size=W3.shape)
    b3[:] = rng.normal(size=b3.shape)
    t, dt, _ = mo.forward(target_p, x, H, O)
    q = mo.refit_output(p, x, t, dt, wd, H, O, ridge=0.0)
    np.testing.assert_allclose(q, target_p, rtol=1e-7, atol=1e-8)
    # and the refit minimises the loss over the output layer: perturbing W3 raises it
    L0, g = mo.loss_and_grad(q, x, t, dt, np.ones(O), wd, H, O)
    assert L0 < 1e-15
