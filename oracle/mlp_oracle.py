"""CPU oracle for the learned selector's network — TEST INFRASTRUCTURE ONLY.

Only tests/ use this module, as the checker of the HIP network kernels
(rvsp_mlp_forward / rvsp_mlp_gradient / rvsp_mlp_train). The product library
never loads it.

The network has no counterpart in the reference (its per-row accumulator
choice, src/backend/rvv/accum_adaptive_f32.c:213-278, is a fixed rule), so
this restatement is pinned by mathematics instead of reference vectors
(tests/test_mlp_oracle.py): central finite differences of the float64 loss
against the analytic gradient, softmax invariances (a constant added to the
allowed logits changes nothing; forbidden classes get no gradient), the
closed-form loss of a zero network, and planted mistakes (a dropped ReLU
mask, a transposed weight) that the finite-difference check must catch.

Shapes and layout follow include/rv_sparse_learned.h: F=16 features, H=64
hidden units (two ReLU layers), K=52 classes; parameters flat, row-major,
W1[H][F], b1[H], W2[H][H], b2[H], W3[K][H], b3[K].
"""
import numpy as np

F, H, K = 17, 64, 56
P = H * F + H + H * H + H + K * H + K


def unpack(params):
    p = np.asarray(params, np.float64)
    assert p.shape == (P,)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + K * H].reshape(K, H); o += K * H
    b3 = p[o:o + K]; o += K
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def mask_bits(mask):
    """(n,) uint64 class masks -> (n, K) bool."""
    m = np.asarray(mask, np.uint64)
    return ((m[:, None] >> np.arange(K, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)


def forward(params, x):
    """Logits and hidden activations, float64."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    z = h2 @ W3.T + b3
    return z, h1, h2


def predict(params, x, mask):
    """Masked argmax (first maximum on ties), -1 when a sample allows none."""
    z, _, _ = forward(params, x)
    mb = mask_bits(mask)
    zz = np.where(mb, z, -np.inf)
    cls = np.argmax(zz, axis=1)
    cls[~mb.any(axis=1)] = -1
    return cls


PLANTS = ("no_relu_mask", "transpose_w2", "softmax_unmasked")


def loss_and_grad(params, x, q, mask, w, _plant=None):
    """sum_i w_i * CE(q_i, softmax over allowed classes of z_i) and its
    gradient with respect to the flat parameters, in float64.

    _plant (tests only) plants one of PLANTS in the gradient, so the tests
    can show that their checks catch it."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    x = np.asarray(x, np.float64)
    q = np.asarray(q, np.float64)
    w = np.asarray(w, np.float64)
    mb = mask_bits(mask)
    z, h1, h2 = forward(params, x)
    zz = np.where(mb, z, -np.inf)
    zmax = np.max(zz, axis=1, keepdims=True)
    zmax = np.where(np.isfinite(zmax), zmax, 0.0)
    e = np.where(mb, np.exp(zz - zmax), 0.0)
    se = e.sum(axis=1, keepdims=True)
    any_ok = mb.any(axis=1)
    lse = zmax + np.log(np.where(se > 0, se, 1.0))
    logp = np.where(mb, z - lse, 0.0)
    loss = float(np.sum(w * np.where(any_ok, -(q * logp * mb).sum(axis=1), 0.0)))
    p = np.where(mb, np.exp(logp), 0.0)
    if _plant == "softmax_unmasked":
        p = np.exp(z - z.max(axis=1, keepdims=True))
        p /= p.sum(axis=1, keepdims=True)
    dz = np.where(mb, w[:, None] * (p - q), 0.0)
    gW3 = dz.T @ h2
    gb3 = dz.sum(axis=0)
    d2 = (dz @ W3) * (h2 > 0 if _plant != "no_relu_mask" else 1.0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ (W2.T if _plant == "transpose_w2" else W2)) * (h1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam(params, x, q, mask, w, epochs, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Full-batch Adam with bias correction; returns (params, losses), where
    losses[e] is the loss before step e."""
    p = np.asarray(params, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for t in range(1, epochs + 1):
        loss, g = loss_and_grad(p, x, q, mask, w)
        losses.append(loss)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mhat = m / (1 - b1 ** t)
        vhat = v / (1 - b2 ** t)
        p = p - lr * mhat / (np.sqrt(vhat) + eps)
    return p, np.array(losses)


def random_problem(n, seed, scale=0.3):
    """Seeded parameters, inputs, soft targets, masks and weights."""
    rng = np.random.default_rng(seed)
    params = rng.normal(0, scale, P)
    x = rng.normal(0, 1, (n, F))
    mb = rng.random((n, K)) < 0.4
    mb[np.arange(n), rng.integers(0, K, n)] = True
    logits = rng.normal(0, 2, (n, K))
    q = np.where(mb, np.exp(logits), 0.0)
    q /= q.sum(axis=1, keepdims=True)
    mask = (mb.astype(np.uint64) << np.arange(K, dtype=np.uint64)[None, :]).sum(
        axis=1, dtype=np.uint64)
    w = rng.random(n) / n
    return params, x, q, mask, w
