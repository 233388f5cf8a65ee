"""Python side of the learned per-bin selector (include/rv_sparse_learned.h).

The library does the work: rvsp_spgemm_autotune records one sample per timed
bin, rvsp_learned_fit trains the network on the device, and
RVSP_SPGEMM_ALGO_HIP_LEARNED descriptors consult it at work_estimation. This
module moves samples in and out as numpy arrays, wraps the raw network
kernels, and scores choices against measured times.
"""
import ctypes as C

import numpy as np

from . import _ffi as F
from ._ffi import check, lib
from .api import DeviceCsr, Descriptor

NF, NH, NK, NP = F.MLP_FEATURES, F.MLP_HIDDEN, F.MLP_CLASSES, F.MLP_PARAMS


def _fp(a):
    return a.ctypes.data_as(F.FP) if a is not None else None


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def class_table():
    """[{kind, group, cap_mult, variant}] for every output class."""
    out = []
    for c in range(NK):
        e = F.LearnedClass()
        check(lib().rvsp_learned_class_info(c, C.byref(e)), "class_info")
        out.append(dict(kind=e.kind, group=e.group, cap_mult=e.cap_mult, variant=e.variant,
                        cap=e.cap,
                        name=f"{F.KIND_NAMES[e.kind]}/g{e.group}x{e.cap_mult}v{e.variant}"
                             + (f"t{e.cap}" if e.cap else "")))
    return out


REUSE = NF - 1  # the network input that marks a numeric-only reuse choice


def map_class_mask():
    """Bit set of the position-map classes (left to the tuner)."""
    m = 0
    for c, e in enumerate(class_table()):
        if e["kind"] == F.KIND_MAP:
            m |= 1 << c
    return m


# ---------------------------------------------------------------------------
# Samples.

def sample_count():
    return int(lib().rvsp_learned_sample_count())


def samples(first=0, count=None):
    """Samples [first, first+count) as a dict of numpy arrays."""
    n = sample_count()
    count = n - first if count is None else count
    phase = np.zeros(count, np.int32)
    bins = np.zeros(count, np.int32)
    fixed = np.zeros(count, np.int32)
    feats = np.zeros((count, NF), np.float32)
    us = np.zeros((count, NK), np.float32)
    s = F.LearnedSample()
    for i in range(count):
        check(lib().rvsp_learned_sample_get(first + i, C.byref(s)), "sample_get")
        phase[i], bins[i], fixed[i] = s.phase, s.bin, s.fixed_class
        feats[i] = np.frombuffer(s.features, np.float32)
        us[i] = np.frombuffer(s.us, np.float32)
    return dict(phase=phase, bin=bins, fixed_class=fixed, features=feats, us=us)


def add_samples(sm):
    s = F.LearnedSample()
    for i in range(len(sm["phase"])):
        s.phase, s.bin, s.fixed_class = int(sm["phase"][i]), int(sm["bin"][i]), int(sm["fixed_class"][i])
        s.n_candidates = int((sm["us"][i] >= 0).sum())
        C.memmove(s.features, _f32(sm["features"][i]).ctypes.data, NF * 4)
        C.memmove(s.us, _f32(sm["us"][i]).ctypes.data, NK * 4)
        check(lib().rvsp_learned_sample_add(C.byref(s)), "sample_add")


def clear_samples():
    lib().rvsp_learned_sample_clear()


def save_samples(path, sm):
    np.savez_compressed(path, **sm)


def load_samples(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def reuse_rows(sm):
    """The numeric samples with timed position-map classes as the library's
    training adds them (rvsp_learned_fit): the reuse input set, and as the
    plain choice the class the plain promotion uses (the fixed search bin's
    lanes, or 64 lanes for a long-row bin, chunked variant)."""
    tab = class_table()
    maps = np.array([t["kind"] == F.KIND_MAP for t in tab], bool)
    keep = (sm["phase"] == 2) & ((sm["us"] >= 0) & maps[None, :]).any(axis=1)
    out = {k: np.array(v)[keep] for k, v in sm.items()}
    out["features"] = out["features"].copy()
    out["features"][:, REUSE] = 1.0
    by = {(t["group"], t["variant"]): c for c, t in enumerate(tab) if t["kind"] == F.KIND_MAP}
    fixed = []
    for c in out["fixed_class"]:
        t = tab[c]
        g = t["group"] if t["kind"] == F.KIND_SEARCH else 64
        fixed.append(by[(g, 2)])
    out["fixed_class"] = np.array(fixed, np.int32)
    return out


def candidate_masks(sm, include_map=False):
    """Per-sample bit sets of the measured classes the network may pick, as
    the library's selection: for rows with the reuse input set, the
    position-map classes only; otherwise no position-map classes (unless
    include_map) and no global dense accumulator for bins whose fixed kind
    is a short-row kernel."""
    tab = class_table()
    timed = sm["us"] >= 0
    maps = np.array([t["kind"] == F.KIND_MAP for t in tab], bool)
    glob = np.array([t["kind"] == F.KIND_GLOBAL for t in tab], bool)
    short = np.array([tab[c]["kind"] in (F.KIND_HASH, F.KIND_SEARCH, F.KIND_DIRECT)
                      for c in sm["fixed_class"]], bool)
    reuse = sm["features"][:, REUSE] > 0.5
    plain = timed & ~(short[:, None] & glob[None, :])
    if not include_map:
        plain = plain & ~maps[None, :]
    ok = np.where(reuse[:, None], timed & maps[None, :], plain)
    return (ok.astype(np.uint64) << np.arange(NK, dtype=np.uint64)[None, :]).sum(
        axis=1, dtype=np.uint64)


def score(sm, cls):
    """Measured time of the chosen classes vs the fixed table's and the best
    measured, summed over the samples (microseconds)."""
    idx = np.arange(len(cls))
    us = sm["us"]
    t_choice = float(us[idx, cls].sum())
    t_fixed = float(us[idx, sm["fixed_class"]].sum())
    best = np.where(us >= 0, us, np.inf).min(axis=1)
    return dict(chosen_us=t_choice, fixed_us=t_fixed, best_us=float(best.sum()),
                gain=t_fixed / t_choice if t_choice > 0 else float("nan"))


# ---------------------------------------------------------------------------
# Model.

def fit(first=0, count=None, epochs=1500, lr=3e-3, seed=1):
    count = sample_count() - first if count is None else count
    loss = C.c_float()
    check(lib().rvsp_learned_fit(first, count, epochs, lr, seed, C.byref(loss)), "learned_fit")
    return float(loss.value)


def predict(features, mask):
    x = _f32(features, (-1, NF))
    m = _u64(mask)
    out = np.zeros(len(x), np.int32)
    check(lib().rvsp_learned_predict(_fp(x), m.ctypes.data_as(F.U64P), len(x),
                                     out.ctypes.data_as(F.I32P)), "learned_predict")
    return out


def save_model(path=None):
    return int(lib().rvsp_learned_model_save(path.encode() if path else None))


def load_model(path=None):
    return int(lib().rvsp_learned_model_load(path.encode() if path else None))


def clear_model():
    lib().rvsp_learned_model_clear()


def model_ready():
    return bool(lib().rvsp_learned_model_ready())


# ---------------------------------------------------------------------------
# Raw network kernels.

def mlp_forward(params, x, mask=None):
    p = _f32(params)
    xx = _f32(x, (-1, NF))
    n = len(xx)
    logits = np.zeros((n, NK), np.float32)
    cls = np.zeros(n, np.int32) if mask is not None else None
    m = _u64(mask) if mask is not None else None
    check(lib().rvsp_mlp_forward(_fp(p), _fp(xx), m.ctypes.data_as(F.U64P) if m is not None else None,
                                 n, _fp(logits), cls.ctypes.data_as(F.I32P) if cls is not None else None),
          "mlp_forward")
    return logits, cls


def mlp_gradient(params, x, q, mask, w):
    p, xx, qq, ww = _f32(params), _f32(x, (-1, NF)), _f32(q, (-1, NK)), _f32(w)
    m = _u64(mask)
    g = np.zeros(NP, np.float32)
    loss = C.c_float()
    check(lib().rvsp_mlp_gradient(_fp(p), _fp(xx), _fp(qq), m.ctypes.data_as(F.U64P), _fp(ww),
                                  len(xx), _fp(g), C.byref(loss)), "mlp_gradient")
    return float(loss.value), g


def mlp_train(params, x, q, mask, w, epochs, lr):
    p = _f32(params).copy()
    xx, qq, ww = _f32(x, (-1, NF)), _f32(q, (-1, NK)), _f32(w)
    m = _u64(mask)
    losses = np.zeros(max(epochs, 1), np.float32)
    check(lib().rvsp_mlp_train(_fp(p), _fp(xx), _fp(qq), m.ctypes.data_as(F.U64P), _fp(ww),
                               len(xx), epochs, lr, _fp(losses)), "mlp_train")
    return p, losses[:epochs]


# ---------------------------------------------------------------------------
# Sample collection.

def autotune_device(A, B=None, reps=3):
    """Runs the autotuner on C = A*B with device-resident operands (which
    records one sample per timed bin); returns the number of new samples."""
    before = sample_count()
    dA = DeviceCsr.upload(A)
    dB = dA if B is None or B is A else DeviceCsr.upload(B)
    d = Descriptor(algo=F.ALGO_HIP_FIXED, memory=F.MEMORY_DEVICE)
    try:
        nnz = d.estimate(dA, dB)
        dC = DeviceCsr(A.rows, dB.cols, nnz)
        try:
            a, b, c = dA.struct(), dB.struct(), dC.struct(nnz)
            check(lib().rvsp_spgemm_autotune(d.handle, C.byref(a), C.byref(b), C.byref(c),
                                             C.c_void_p(d.ws.ctypes.data), reps), "autotune")
        finally:
            dC.free()
    finally:
        d.close()
        dA.free()
        if dB is not dA:
            dB.free()
    return sample_count() - before
