"""Pins the CPU oracle of the selector network (oracle/mlp_oracle.py) by
mathematics: central finite differences of the float64 loss, softmax
invariances, the closed-form loss of a zero network and Adam's first step.
Each planted mistake in the oracle must fail the finite-difference check."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def fd_check(loss_fn, params, grad, idx, h=1e-6):
    """Largest relative disagreement between grad[idx] and central finite
    differences of loss_fn at params."""
    worst = 0.0
    scale = max(1e-12, np.max(np.abs(grad)))
    for i in idx:
        p1 = params.copy(); p1[i] += h
        p0 = params.copy(); p0[i] -= h
        fd = (loss_fn(p1) - loss_fn(p0)) / (2 * h)
        worst = max(worst, abs(fd - grad[i]) / scale)
    return worst


def sample_idx(seed, n=240):
    rng = np.random.default_rng(seed)
    # Every parameter block is represented.
    blocks = [(0, M.H * M.F), (M.H * M.F, M.H), (M.H * M.F + M.H, M.H * M.H),
              (M.H * M.F + M.H + M.H * M.H, M.H),
              (M.H * M.F + 2 * M.H + M.H * M.H, M.K * M.H), (M.P - M.K, M.K)]
    idx = []
    for off, size in blocks:
        idx += list(off + rng.choice(size, min(size, n // len(blocks)), replace=False))
    return idx


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    params, x, q, mask, w = M.random_problem(24, seed)
    loss, g = M.loss_and_grad(params, x, q, mask, w)
    assert loss > 0
    err = fd_check(lambda p: M.loss_and_grad(p, x, q, mask, w)[0], params, g, sample_idx(seed))
    assert err < 1e-6, err


@pytest.mark.parametrize("plant", M.PLANTS)
def test_planted_mistakes_fail_finite_differences(plant):
    params, x, q, mask, w = M.random_problem(24, 7)
    _, g = M.loss_and_grad(params, x, q, mask, w, _plant=plant)
    err = fd_check(lambda p: M.loss_and_grad(p, x, q, mask, w)[0], params, g, sample_idx(7))
    assert err > 1e-3, f"planted mistake {plant} not caught ({err})"


def test_softmax_invariances():
    params, x, q, mask, w = M.random_problem(16, 3)
    loss, g = M.loss_and_grad(params, x, q, mask, w)
    # A constant added to every logit changes nothing.
    p2 = params.copy()
    p2[M.P - M.K:] += 5.0
    assert abs(M.loss_and_grad(p2, x, q, mask, w)[0] - loss) < 1e-9 * max(1, loss)
    # ... so the output-bias gradient sums to zero.
    assert abs(g[M.P - M.K:].sum()) < 1e-12 + 1e-9 * np.abs(g).max()
    # Classes no sample allows get no gradient.
    mb = M.mask_bits(mask)
    never = ~mb.any(axis=0)
    mask2 = mask & np.uint64((1 << 40) - 1)  # forbid classes 40..51 everywhere
    q2 = np.where(M.mask_bits(mask2), q, 0.0)
    rows = q2.sum(axis=1) > 0
    q2[rows] /= q2[rows].sum(axis=1, keepdims=True)
    _, g2 = M.loss_and_grad(params, x, q2, mask2, w)
    assert np.all(g2[M.P - M.K + 40:] == 0)
    W3g = g2[M.P - M.K - M.K * M.H:M.P - M.K].reshape(M.K, M.H)
    assert np.all(W3g[40:] == 0)
    del never


def test_zero_network_closed_form():
    params, x, q, mask, w = M.random_problem(32, 4)
    zero = np.zeros(M.P)
    loss, _ = M.loss_and_grad(zero, x, q, mask, w)
    n_allowed = M.mask_bits(mask).sum(axis=1)
    assert abs(loss - float(np.sum(w * np.log(n_allowed)))) < 1e-12


def test_predict_is_masked_argmax():
    params, x, q, mask, w = M.random_problem(50, 5)
    z, _, _ = M.forward(params, x)
    cls = M.predict(params, x, mask)
    mb = M.mask_bits(mask)
    for i in range(50):
        allowed = np.flatnonzero(mb[i])
        assert cls[i] == allowed[np.argmax(z[i, allowed])]
    none = np.zeros(3, np.uint64)
    assert list(M.predict(params, x[:3], none)) == [-1, -1, -1]


def test_adam_first_step_is_signed_lr():
    params, x, q, mask, w = M.random_problem(16, 6)
    _, g = M.loss_and_grad(params, x, q, mask, w)
    p1, losses = M.adam(params, x, q, mask, w, 1, 1e-3)
    step = params - p1
    big = np.abs(g) > 1e-4  # |g| / (|g| + 1e-8) within 1e-4 of 1
    assert np.allclose(step[big], 1e-3 * np.sign(g[big]), rtol=1e-3)
    assert np.all(step[g == 0] == 0)
    assert losses[0] == pytest.approx(M.loss_and_grad(params, x, q, mask, w)[0])


def test_adam_reduces_loss():
    params, x, q, mask, w = M.random_problem(40, 8, scale=0.1)
    _, losses = M.adam(params, x, q, mask, w, 60, 3e-3)
    assert losses[-1] < 0.8 * losses[0]
