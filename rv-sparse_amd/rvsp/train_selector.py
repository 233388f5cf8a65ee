"""Builds the learned selector's default model on the current GPU.

    python -m rvsp.train_selector [--model PATH] [--samples PATH] [--reps N]

Runs the autotuner (which times every candidate kernel shape of every
non-empty bin and records one sample per bin) on a training set of
products, trains the network on the device from those samples and writes
the model (by default learned_model.txt next to librvsparse_hip.so, where
RVSP_SPGEMM_ALGO_HIP_LEARNED descriptors find it) and the samples (.npz).

Training set: R-MAT (a,b,c,d = .57,.19,.19,.05, edge factor 16) at scales
14-17 with seeds 1..8 and scale 18 with seeds 1..3 (seed 0 is kept for
evaluation), plus lognormal-row products (up to 2^22 rows, the size of a
stacked batch's bins), uniform-k products, 5-point
Laplacians and banded rows with a few dense rows at three sizes, so that
every bin of every phase is seen with short and long B rows.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import _ffi as F
from . import gen, learned


def training_set():
    """(name, make) with make() -> (A, B); B None means A*A."""
    for scale in (14, 15, 16, 17):
        for seed in range(1, 9):
            yield f"rmat s{scale} seed{seed}", (lambda s=scale, k=seed: (gen.rmat(s, seed=k), None))
    for seed in range(1, 4):
        yield f"rmat s18 seed{seed}", (lambda k=seed: (gen.rmat(18, seed=k), None))
    for n in (4096, 32768, 131072, 1 << 20, 1 << 22):
        for seed in (1, 2):
            yield f"lognormal n{n} seed{seed}", (
                lambda n=n, k=seed: (gen.lognormal_rows(n, seed=k),
                                     gen.lognormal_rows(n, seed=k + 100)))
    for k in (4, 16, 48):
        yield f"uniform n65536 k{k}", (lambda k=k: (gen.uniform_k(65536, k, seed=k), None))
    for g in (256, 1024):
        yield f"laplacian g{g}", (lambda g=g: (gen.laplacian_2d(g), None))
    yield "banded 50k", (lambda: gen.banded_mixed(rows=50_000, inner=20_000, cols=50_000,
                                                  seed=1))
    for rows, seed in ((100_000, 2), (250_000, 3)):
        yield f"banded {rows // 1000}k", (
            lambda r=rows, k=seed: gen.banded_mixed(rows=r, inner=int(r * 0.4), cols=r, seed=k))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[1])
    ap.add_argument("--model", default=None, help="model file (default: next to the library)")
    ap.add_argument("--samples", default=os.path.join(os.path.dirname(F.LIB_PATH),
                                                      "learned_samples.npz"))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--only-fit", action="store_true",
                    help="train from --samples instead of collecting")
    args = ap.parse_args(argv)
    learned.clear_samples()
    if args.only_fit:
        learned.add_samples(learned.load_samples(args.samples))
    else:
        for name, make in training_set():
            t0 = time.time()
            A, B = make()
            new = learned.autotune_device(A, B, reps=args.reps)
            print(f"{name}: {A.rows} rows, {A.nnz} nnz, {new} samples, "
                  f"{time.time() - t0:.1f} s", flush=True)
        learned.save_samples(args.samples, learned.samples())
    n = learned.sample_count()
    loss = learned.fit(0, n, epochs=args.epochs, lr=args.lr, seed=args.seed)
    sm = learned.samples()
    cls = learned.predict(sm["features"], learned.candidate_masks(sm))
    sc = learned.score(sm, cls)
    print(f"trained on {n} samples: loss {loss:.4f}; training-set time fixed "
          f"{sc['fixed_us']:.0f} us, learned {sc['chosen_us']:.0f} us, best {sc['best_us']:.0f} us",
          flush=True)
    if learned.save_model(args.model) != 0:
        print("could not write the model", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
