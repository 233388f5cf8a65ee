"""The learned selector on the device: the network kernels against the
float64 oracle (oracle/mlp_oracle.py), training, and the selector's effect:
unchanged output bits and measured time against the fixed table."""
import numpy as np
import pytest

import rvsp
from oracle import mlp_oracle as M
from oracle import oracle
from rvsp import _ffi as F
from rvsp import gen, learned

pytestmark = pytest.mark.gpu


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [1, 33, 100])
def test_forward_matches_oracle(gpu, n):
    params, x, q, mask, w = M.random_problem(n, 11 + n)
    if n > 2:
        mask[2] = 0  # a sample that allows no class
    logits, cls = learned.mlp_forward(params, x, mask)
    z, _, _ = M.forward(f32(params), f32(x))
    # float32 device arithmetic against float64: 1e-4 of the logit scale.
    tol = 1e-4 * np.abs(z).max() + 1e-5
    assert np.abs(logits - z).max() <= tol
    want = M.predict(f32(params), f32(x), mask)
    zz = np.where(M.mask_bits(mask), z, -np.inf)
    srt = np.sort(zz, axis=1)
    with np.errstate(invalid="ignore"):
        clear = (srt[:, -1] - srt[:, -2] > 10 * tol) | ~np.isfinite(srt[:, -2])
    assert np.array_equal(cls[clear], want[clear])
    if n > 2:
        assert cls[2] == -1


@pytest.mark.parametrize("n", [1, 32, 100])
def test_gradient_matches_oracle(gpu, n):
    params, x, q, mask, w = M.random_problem(n, 21 + n)
    if n > 2:
        mask[1] = 0
        q[1] = 0
    loss, g = learned.mlp_gradient(params, x, q, mask, w)
    lr_, gr = M.loss_and_grad(f32(params), f32(x), f32(q), mask, f32(w))
    # Tolerances: loss to 1e-5 relative; every gradient entry to 1e-4 of the
    # largest entry of its parameter block.
    assert abs(loss - lr_) <= 1e-5 * abs(lr_) + 1e-7
    o = 0
    for size in (M.H * M.F, M.H, M.H * M.H, M.H, M.K * M.H, M.K):
        blk, ref = g[o:o + size], gr[o:o + size]
        assert np.abs(blk - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-9, o
        o += size


def test_training_steps_match_oracle_adam(gpu):
    params, x, q, mask, w = M.random_problem(64, 5)
    epochs, lr = 5, 1e-3
    pg, lg = learned.mlp_train(params, x, q, mask, w, epochs, lr)
    pr, lr_ = M.adam(f32(params), f32(x), f32(q), mask, f32(w), epochs, lr)
    assert np.allclose(lg, lr_, rtol=1e-4, atol=1e-7)
    d = np.abs(pg - pr)
    # An Adam step moves a parameter by at most about lr; entries whose
    # gradient is near float32 noise may differ by up to that, the rest
    # agree to 1e-3 of a step.
    assert d.max() <= 2 * lr * epochs
    assert np.quantile(d, 0.99) <= 1e-3 * lr * epochs


def test_training_reduces_loss_toward_teacher(gpu):
    rng = np.random.default_rng(3)
    teacher = rng.normal(0, 0.4, M.P)
    student = rng.normal(0, 0.1, M.P)
    n = 512
    x = rng.normal(0, 1, (n, M.F))
    mb = rng.random((n, M.K)) < 0.5
    mb[np.arange(n), rng.integers(0, M.K, n)] = True
    mask = (mb.astype(np.uint64) << np.arange(M.K, dtype=np.uint64)[None, :]).sum(
        axis=1, dtype=np.uint64)
    z, _, _ = M.forward(teacher, x)
    e = np.where(mb, np.exp(z - np.where(mb, z, -np.inf).max(axis=1, keepdims=True)), 0)
    q = e / e.sum(axis=1, keepdims=True)
    w = np.full(n, 1.0 / n)
    entropy = float(-(w[:, None] * np.where(q > 0, q * np.log(np.where(q > 0, q, 1)), 0)).sum())
    _, losses = learned.mlp_train(student, x, q, mask, w, 400, 3e-3)
    assert losses[-1] < losses[0]
    # The excess over the teacher's entropy (the KL divergence) halves.
    assert losses[-1] - entropy < 0.5 * (losses[0] - entropy)


def test_learned_selector_keeps_bits_and_gains_over_fixed(gpu):
    """Trains on timings the tuner records for R-MAT products of seeds 1..8
    and scores the network's choices on seed 0, which it has not seen,
    against the fixed table, both priced with seed 0's own measured times."""
    learned.clear_samples()
    for seed in range(1, 9):
        assert learned.autotune_device(gen.rmat(15, seed=seed), reps=3) > 0
    n_train = learned.sample_count()
    A = gen.rmat(15, seed=0)
    assert learned.autotune_device(A, reps=3) > 0
    loss = learned.fit(0, n_train, epochs=1500, lr=3e-3, seed=1)
    assert np.isfinite(loss)
    ev = learned.samples(n_train)
    cls = learned.predict(ev["features"], learned.candidate_masks(ev))
    assert np.all(cls >= 0)
    sc = learned.score(ev, cls)
    print(f"\nlearned selector on unseen R-MAT 2^15 seed 0: fixed {sc['fixed_us']:.1f} us, "
          f"learned {sc['chosen_us']:.1f} us, best measured {sc['best_us']:.1f} us, "
          f"gain {sc['gain']:.3f}x over {len(cls)} bins")
    assert sc["chosen_us"] < sc["fixed_us"]
    # Reuse choices (the first numeric-only compute re-shapes the bins moved
    # to the position map): against the plain promotion's map shape.
    ev_r = learned.reuse_rows(ev)
    if len(ev_r["phase"]):
        cls_r = learned.predict(ev_r["features"], learned.candidate_masks(ev_r))
        assert np.all(cls_r >= 0)
        sr = learned.score(ev_r, cls_r)
        print(f"reuse choices over {len(cls_r)} bins: plain promotion {sr['fixed_us']:.1f} us, "
              f"learned {sr['chosen_us']:.1f} us, best {sr['best_us']:.1f} us")
        assert sr["chosen_us"] <= sr["fixed_us"]
    # The selector changes kernel shapes only: a HIP_LEARNED product is
    # bit-identical to the oracle.
    want = oracle.spgemm(A, A)
    got, plan = rvsp.spgemm_device(A, A, algo=F.ALGO_HIP_LEARNED)
    for g, w_ in zip((got.row_ptr, got.col_idx, got.values), want):
        assert g.tobytes() == w_.tobytes()
    learned.clear_model()
    learned.clear_samples()


def test_learned_without_model_is_the_fixed_plan(gpu):
    learned.clear_model()
    A = gen.uniform_k(2048, 12, seed=9)
    got, plan = rvsp.spgemm_device(A, A, algo=F.ALGO_HIP_LEARNED)
    _, plan_fixed = rvsp.spgemm_device(A, A, algo=F.ALGO_HIP_FIXED)
    assert plan == plan_fixed
    want = oracle.spgemm(A, A)
    assert got.values.tobytes() == want[2].tobytes()
