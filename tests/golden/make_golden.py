"""Records golden vectors from the reference library built from its own
sources (oracle/_ref/librvsp_ref.so, `make -C oracle ref`).

For every case in tests/cases.py GOLDEN_CASES it stores the SHA-256 digest of
the reference output (row_ptr, col_idx, values bytes), nnz(C), a digest of
the per-row op counts, and a few sampled rows with exact float bits. Run in
the container that holds the reference tree:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import cases  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/librvsp_ref.so missing: run `make -C oracle ref` first")
    out = {"source": "reference raw scalar kernel rvsp_spgemm_scalar_f32 "
                     "(gustavson_core_f32.inc:66-194) built from its own sources",
           "cases": {}}
    for name in cases.GOLDEN_CASES:
        A, B = cases.build(name)
        rp, ci, val, ops = oracle.ref_spgemm(A, B, want_ops=True)
        rng = np.random.Generator(np.random.PCG64(1234))
        nonempty = np.nonzero(np.diff(rp))[0]
        pick = sorted(set(rng.choice(nonempty, size=min(4, nonempty.size), replace=False).tolist()))
        samples = []
        for r in pick:
            s, e = int(rp[r]), int(rp[r + 1])
            samples.append({"row": int(r), "col_idx": ci[s:e][:64].tolist(),
                            "value_bits": val[s:e][:64].view(np.uint32).tolist()})
        out["cases"][name] = {
            "a": [A.rows, A.cols, A.nnz], "b": [B.rows, B.cols, B.nnz],
            "nnz_c": int(rp[-1]), "products": int(ops.sum()),
            "digest": oracle.digest(rp, ci, val),
            "ops_digest": oracle.digest(ops.astype(np.int32), np.zeros(0, np.int32),
                                        np.zeros(0, np.float32)) if ops.max() < 2**31 else None,
            "samples": samples,
        }
        print(f"{name:28s} nnz_c={int(rp[-1]):>10d} {out['cases'][name]['digest'][:16]}")
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
