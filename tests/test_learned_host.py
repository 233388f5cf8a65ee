"""Host-side contract of the learned selector's C-ABI (include/
rv_sparse_learned.h): the class table, the sample store, model files and
argument checks. Nothing here launches a kernel."""
import ctypes as C

import numpy as np
import pytest

from rvsp import _ffi as F
from rvsp import learned

L = F.lib()


def test_class_table_layout():
    tab = learned.class_table()
    assert len(tab) == F.MLP_CLASSES
    kinds = [t["kind"] for t in tab]
    assert kinds.count(F.KIND_HASH) == 14
    assert kinds.count(F.KIND_SEARCH) == 7
    assert kinds.count(F.KIND_MAP) == 21
    assert kinds.count(F.KIND_DIRECT) == 7
    assert kinds.count(F.KIND_BITMAP) == kinds.count(F.KIND_GLOBAL) == 1
    # Long rows: workgroup tiles of 6144, 4096 and 3072 positions, wave
    # tiles of 512 and 1024.
    assert sorted((t["variant"], t["cap"]) for t in tab if t["kind"] == F.KIND_BLOCK) == \
        [(0, 3072), (0, 4096), (0, 6144), (1, 512), (1, 1024)]
    assert all(t["cap"] == 0 for t in tab if t["kind"] != F.KIND_BLOCK)
    # Every lane count of a sub-wave kind appears once per (kind, table size, variant).
    seen = set()
    for t in tab:
        key = (t["kind"], t["group"], t["cap_mult"], t["variant"], t["cap"])
        assert key not in seen
        seen.add(key)
    assert {t["group"] for t in tab if t["kind"] == F.KIND_HASH} == {1, 2, 4, 8, 16, 32, 64}
    assert {t["cap_mult"] for t in tab if t["kind"] == F.KIND_HASH} == {1, 2}
    e = F.LearnedClass()
    assert L.rvsp_learned_class_info(F.MLP_CLASSES, C.byref(e)) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_learned_class_info(-1, C.byref(e)) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_learned_class_info(0, None) == F.ERROR_NULL_POINTER


def test_mlp_dimensions_match_oracle():
    from oracle import mlp_oracle as M
    assert (F.MLP_FEATURES, F.MLP_HIDDEN, F.MLP_CLASSES, F.MLP_PARAMS) == (M.F, M.H, M.K, M.P)


def test_sample_store_round_trip(tmp_path):
    learned.clear_samples()
    rng = np.random.default_rng(0)
    sm = dict(phase=np.array([0, 2, 1], np.int32), bin=np.array([3, 5, 1], np.int32),
              fixed_class=np.array([2, 17, 49], np.int32),
              features=rng.random((3, F.MLP_FEATURES)).astype(np.float32),
              us=np.where(rng.random((3, F.MLP_CLASSES)) < 0.5, -1.0,
                          rng.random((3, F.MLP_CLASSES)) * 100).astype(np.float32))
    learned.add_samples(sm)
    assert learned.sample_count() == 3
    got = learned.samples()
    for k in sm:
        assert np.array_equal(got[k], sm[k]), k
    path = str(tmp_path / "s.npz")
    learned.save_samples(path, got)
    back = learned.load_samples(path)
    assert all(np.array_equal(back[k], sm[k]) for k in sm)
    s = F.LearnedSample()
    assert L.rvsp_learned_sample_get(3, C.byref(s)) == F.ERROR_INVALID_ARGUMENT
    s.phase = 3
    assert L.rvsp_learned_sample_add(C.byref(s)) == F.ERROR_INVALID_ARGUMENT
    learned.clear_samples()
    assert learned.sample_count() == 0


def test_fit_rejects_bad_slices():
    learned.clear_samples()
    assert L.rvsp_learned_fit(0, 1, 10, 1e-3, 1, None) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_learned_fit(0, 0, 10, 1e-3, 1, None) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_learned_fit(0, 1, 0, 1e-3, 1, None) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_learned_fit(0, 1, 10, 0.0, 1, None) == F.ERROR_INVALID_ARGUMENT


def write_model(path, params, mu, sigma):
    with open(path, "w") as f:
        f.write(f"rvsp-mlp {F.MLP_FEATURES} {F.MLP_HIDDEN} {F.MLP_CLASSES}\n")
        f.write("mu " + " ".join(f"{v:.9g}" for v in mu) + "\n")
        f.write("sigma " + " ".join(f"{v:.9g}" for v in sigma) + "\n")
        f.write("params\n" + " ".join(f"{v:.9g}" for v in params) + "\n")


def test_model_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    params = rng.normal(0, 0.1, F.MLP_PARAMS).astype(np.float32)
    mu = rng.random(F.MLP_FEATURES).astype(np.float32)
    sigma = (1 + rng.random(F.MLP_FEATURES)).astype(np.float32)
    p1 = str(tmp_path / "m1.txt")
    write_model(p1, params, mu, sigma)
    learned.clear_model()
    assert not learned.model_ready()
    assert learned.save_model(str(tmp_path / "none.txt")) == -1
    assert learned.load_model(p1) == 0
    assert learned.model_ready()
    p2 = str(tmp_path / "m2.txt")
    assert learned.save_model(p2) == 0
    toks1 = open(p1).read().split()
    toks2 = open(p2).read().split()
    assert len(toks1) == len(toks2)
    assert all(float(a) == float(b) for a, b in zip(toks1[4:], toks2[4:])
               if a not in ("mu", "sigma", "params"))
    assert np.array_equal(np.array([float(t) for t in toks2[-F.MLP_PARAMS:]], np.float32), params)
    # A truncated or mis-sized file is refused and leaves the model as it was.
    bad = str(tmp_path / "bad.txt")
    with open(bad, "w") as f:
        f.write("rvsp-mlp 16 32 52\n")
    assert learned.load_model(bad) == -1
    assert learned.model_ready()
    learned.clear_model()
    assert not learned.model_ready()


def test_predict_without_model_is_refused():
    learned.clear_model()
    x = np.zeros((1, F.MLP_FEATURES), np.float32)
    m = np.ones(1, np.uint64)
    out = np.zeros(1, np.int32)
    st = L.rvsp_learned_predict(x.ctypes.data_as(F.FP), m.ctypes.data_as(F.U64P), 1,
                                out.ctypes.data_as(F.I32P))
    assert st == F.ERROR_INVALID_ARGUMENT


def test_network_entry_points_check_arguments():
    p = np.zeros(F.MLP_PARAMS, np.float32)
    x = np.zeros((1, F.MLP_FEATURES), np.float32)
    assert L.rvsp_mlp_forward(p.ctypes.data_as(F.FP), x.ctypes.data_as(F.FP), None, 0,
                              None, None) == F.SUCCESS
    assert L.rvsp_mlp_forward(p.ctypes.data_as(F.FP), x.ctypes.data_as(F.FP), None, -1,
                              None, None) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_mlp_gradient(None, None, None, None, None, 0, None, None) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_mlp_gradient(None, None, None, None, None, 1, None, None) == F.ERROR_NULL_POINTER
    assert L.rvsp_mlp_train(None, None, None, None, None, 1, -1, 1e-3, None) == F.ERROR_INVALID_ARGUMENT
    assert L.rvsp_mlp_train(None, None, None, None, None, 1, 1, 1e-3, None) == F.ERROR_NULL_POINTER
