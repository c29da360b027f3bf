"""ctypes binding of the surrogate C ABI (include/hwtsurr.h), in the same
library as the tree pricer.  No CPU fallback: every call goes to the GPU.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import capi
from .capi import OptionsC, YieldC, _check, lib

NPARAMS = 199937
HIDDEN = 256
INPUTS = 8
MAX_MEMBERS = 4
RESIDUAL_SCALE = 0.01
OFFSETS = {"W1": 0, "b1": 2048, "W2": 2304, "b2": 67840, "W3": 68096, "b3": 133632,
           "W4": 133888, "b4": 199424, "W5": 199680, "b5": 199936}
SHAPES = {"W1": (256, 8), "b1": (256,), "W2": (256, 256), "b2": (256,), "W3": (256, 256), "b3": (256,),
          "W4": (256, 256), "b4": (256,), "W5": (1, 256), "b5": (1,)}

EXPORTS = (
    "hwt_surr_config_init", "hwt_surr_create", "hwt_surr_destroy", "hwt_surr_set_training_data",
    "hwt_surr_train", "hwt_surr_predict", "hwt_surr_price", "hwt_surr_get_params",
    "hwt_surr_set_params", "hwt_surr_eval_rows", "hwt_surr_get_inputs", "hwt_surr_adam_step",
    "hwt_surr_save", "hwt_surr_load",
)


class SurrConfigC(C.Structure):
    _fields_ = [("members", C.c_int32), ("batch", C.c_int32), ("lr_max", C.c_float), ("lr_min", C.c_float),
                ("warmup_frac", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("spread_threshold", C.c_float), ("reserved0", C.c_int32), ("seed", C.c_uint64),
                ("reserved", C.c_int32 * 6)]


class SurrStatsC(C.Structure):
    _fields_ = [("ms", C.c_double), ("loss", C.c_double), ("steps", C.c_int64), ("fallback", C.c_int32),
                ("out_of_box", C.c_int32), ("surrogate_ms", C.c_double), ("tree_ms", C.c_double),
                ("reserved", C.c_int32 * 4)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        P = C.c_void_p
        L.hwt_surr_config_init.argtypes = [C.POINTER(SurrConfigC)]
        L.hwt_surr_config_init.restype = None
        L.hwt_surr_create.argtypes = [C.POINTER(P), C.c_int32, C.c_int32, C.c_int32, C.POINTER(SurrConfigC)]
        L.hwt_surr_destroy.argtypes = [P]
        L.hwt_surr_destroy.restype = None
        L.hwt_surr_set_training_data.argtypes = [P, C.POINTER(OptionsC), C.POINTER(YieldC), P]
        L.hwt_surr_train.argtypes = [P, C.c_int32, C.c_int32, C.POINTER(SurrStatsC)]
        L.hwt_surr_predict.argtypes = [P, C.POINTER(OptionsC), C.POINTER(YieldC), P, P, P]
        L.hwt_surr_price.argtypes = [P, P, C.POINTER(OptionsC), C.POINTER(YieldC), P, C.POINTER(SurrStatsC)]
        L.hwt_surr_get_params.argtypes = [P, P]
        L.hwt_surr_set_params.argtypes = [P, P]
        L.hwt_surr_eval_rows.argtypes = [P, C.c_int32, P, C.c_int32, P, P, P]
        L.hwt_surr_get_inputs.argtypes = [P, P, C.c_int32, P, P]
        L.hwt_surr_adam_step.argtypes = [P, P, C.c_float]
        L.hwt_surr_save.argtypes = [P, C.c_char_p]
        L.hwt_surr_load.argtypes = [P, C.c_char_p]
        _bound = True
    return L


def config(**kw) -> SurrConfigC:
    c = SurrConfigC()
    _lib().hwt_surr_config_init(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _dev_opts(dev: dict, n: int | None = None) -> OptionsC:
    n = int(dev["n"] if n is None else n)
    return OptionsC(n, *[dev[f].value for f in capi.Plan.FIELDS])


def _dev_yield(dev: dict) -> YieldC:
    return YieldC(int(dev["ny"]), dev["rates"].value, dev["days"].value)


def split_params(flat: np.ndarray) -> dict:
    """One member's flat parameter vector -> named tensors (views)."""
    return {k: flat[o:o + int(np.prod(SHAPES[k]))].reshape(SHAPES[k]) for k, o in OFFSETS.items()}


class Surrogate:
    """An ensemble surrogate on the GPU (hwt_surrogate)."""

    def __init__(self, train_capacity: int, infer_capacity: int, yield_capacity: int,
                 cfg: SurrConfigC | None = None):
        self.cfg = cfg if cfg is not None else config()
        self._s = C.c_void_p()
        _check(_lib().hwt_surr_create(C.byref(self._s), int(train_capacity), int(infer_capacity),
                                      int(yield_capacity), C.byref(self.cfg)))

    @property
    def members(self) -> int:
        return int(self.cfg.members)

    def set_training_data(self, dev: dict, prices):
        """dev: device options/yield (capi.to_device); prices: device FP64 tree prices."""
        _check(_lib().hwt_surr_set_training_data(self._s, C.byref(_dev_opts(dev)), C.byref(_dev_yield(dev)),
                                                 prices.value))

    def train(self, epochs: int, total_epochs: int) -> dict:
        st = SurrStatsC()
        _check(_lib().hwt_surr_train(self._s, int(epochs), int(total_epochs), C.byref(st)))
        return st.as_dict()

    def predict(self, dev: dict):
        """Ensemble-mean price, spread and in-box flag per option (host arrays)."""
        from .devmem import DeviceArray
        n = int(dev["n"])
        mean, spread = DeviceArray.empty((n,), np.float64), DeviceArray.empty((n,), np.float64)
        inbox = DeviceArray.empty((n,), np.int8)
        _check(_lib().hwt_surr_predict(self._s, C.byref(_dev_opts(dev)), C.byref(_dev_yield(dev)),
                                       mean.value, spread.value, inbox.value))
        return mean.to_host(), spread.to_host(), inbox.to_host().astype(bool)

    def price(self, tree: capi.Plan, dev: dict, out=None) -> dict:
        """Surrogate + tree fallback into out (default dev['out'])."""
        st = SurrStatsC()
        out = dev["out"] if out is None else out
        _check(_lib().hwt_surr_price(self._s, tree._p, C.byref(_dev_opts(dev)), C.byref(_dev_yield(dev)),
                                     out.value, C.byref(st)))
        return st.as_dict()

    def get_params(self) -> np.ndarray:
        P = np.empty((self.members, NPARAMS), np.float32)
        _check(_lib().hwt_surr_get_params(self._s, P.ctypes.data))
        return P

    def set_params(self, P: np.ndarray):
        P = np.ascontiguousarray(P, dtype=np.float32).reshape(self.members, NPARAMS)
        _check(_lib().hwt_surr_set_params(self._s, P.ctypes.data))

    def eval_rows(self, member: int, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        out = np.empty(rows.size, np.float32)
        grad = np.empty(NPARAMS, np.float32)
        loss = C.c_float()
        _check(_lib().hwt_surr_eval_rows(self._s, int(member), rows.ctypes.data, rows.size, out.ctypes.data,
                                         C.byref(loss), grad.ctypes.data))
        return out, float(loss.value), grad

    def get_inputs(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        x = np.empty((rows.size, INPUTS), np.float32)
        t = np.empty(rows.size, np.float32)
        _check(_lib().hwt_surr_get_inputs(self._s, rows.ctypes.data, rows.size, x.ctypes.data, t.ctypes.data))
        return x, t

    def adam_step(self, grad: np.ndarray, lr: float):
        g = np.ascontiguousarray(grad, dtype=np.float32).reshape(self.members, NPARAMS)
        _check(_lib().hwt_surr_adam_step(self._s, g.ctypes.data, float(lr)))

    def save(self, path):
        _check(_lib().hwt_surr_save(self._s, str(path).encode()))

    def load(self, path):
        _check(_lib().hwt_surr_load(self._s, str(path).encode()))

    def close(self):
        if self._s:
            _lib().hwt_surr_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def training_set(seed: int, n: int, y, tree_plan_capacity: int | None = None):
    """Box samples priced by the FP64 tree on the GPU: (options, device
    buffers, device prices).  The surrogate's training data generator."""
    from . import datasets
    opts = datasets.box(seed, n)
    dev = capi.to_device(opts, y)
    plan = capi.Plan(tree_plan_capacity or n, y.N)
    plan.execute(dev)
    plan.sync()
    plan.close()
    return opts, dev, dev["out"]


def fit(y, n_train: int = 2_000_000, epochs: int = 40, seed: int = 11, infer_capacity: int = 1 << 20,
        cfg: SurrConfigC | None = None, log=None):
    """Trains an ensemble on `n_train` box samples priced by the FP64 tree:
    (surrogate, record).  The record holds the data-generation time, the
    per-epoch losses and the total training time (GPU, HIP events)."""
    import time
    from . import devmem
    t0 = time.perf_counter()
    opts, dev, prices = training_set(seed, n_train, y)
    devmem.synchronize()
    t_data = time.perf_counter() - t0
    s = Surrogate(n_train, infer_capacity, y.N, cfg)
    s.set_training_data(dev, prices)
    losses, train_ms = [], 0.0
    for ep in range(epochs):
        st = s.train(1, epochs)
        losses.append(st["loss"])
        train_ms += st["ms"]
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {st['loss']:.4e}, {st['ms']:.0f} ms")
    for k in list(dev):
        if hasattr(dev[k], "free"):
            dev[k].free()
    rec = {"samples": n_train, "epochs": epochs, "members": s.members, "batch": int(s.cfg.batch),
           "data_s": t_data, "train_s": train_ms * 1e-3, "steps": epochs * (n_train // int(s.cfg.batch)),
           "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None}
    return s, rec


def calibrate_threshold(s: Surrogate, y, n: int = 1_000_000, seed: int = 77, target: float = 1e-3,
                        margin: float = 0.5) -> dict:
    """The largest ensemble-spread threshold at which every option the
    surrogate would price in a held-out sample (box_mix, its own seed) stays
    within margin * target of the FP64 tree.  In-box options are sorted by
    spread; the threshold is the spread of the last option before the running
    maximum error first exceeds the bound."""
    from . import datasets
    opts = datasets.box_mix(seed, n)
    dev = capi.to_device(opts, y)
    plan = capi.Plan(n, y.N)
    plan.execute(dev)
    plan.sync()
    plan.close()
    tree = dev["out"].to_host()
    mean, spread, inbox = s.predict(dev)
    for k in list(dev):
        if hasattr(dev[k], "free"):
            dev[k].free()
    thr, accepted = threshold_from(tree, mean, spread, inbox, margin * target)
    return {"threshold": thr, "calibration_options": n, "calibration_seed": seed,
            "accepted_frac": accepted / n, "bound": margin * target}


def threshold_from(tree, mean, spread, inbox, bound: float):
    """(threshold, accepted count): the largest spread threshold at which every
    in-box option with spread <= threshold has |mean - tree| <= bound.  Ties
    at the threshold must all pass; with none passing, threshold 0 (an
    option's spread can be exactly 0 only if all members agree to the bit)."""
    idx = np.nonzero(inbox)[0]
    order = idx[np.argsort(spread[idx], kind="stable")]
    run = np.maximum.accumulate(np.abs(mean[order] - tree[order]))
    ok = np.nonzero(run <= bound)[0]
    k = int(ok[-1]) if ok.size and ok[0] == 0 else -1
    while k >= 0 and k + 1 < order.size and spread[order[k + 1]] == spread[order[k]]:
        k -= 1
    return (float(spread[order[k]]) if k >= 0 else 0.0), k + 1


def with_threshold(s: Surrogate, threshold: float, infer_capacity: int, yield_capacity: int) -> Surrogate:
    """A copy of the ensemble (same parameters) gating at another spread threshold."""
    cfg = config(members=s.members, spread_threshold=threshold)
    t = Surrogate(0, infer_capacity, yield_capacity, cfg)
    t.set_params(s.get_params())
    return t


def fallback_plan(n: int, yield_capacity: int, wide_per_cu: int = 0, narrow_per_cu: int = 0) -> capi.Plan:
    """The tree plan hwt_surr_price re-prices with (no graph; optionally fewer
    resident workgroups per CU, leaving room for the inference beside it)."""
    cfg = capi.config(flags=capi.FLAG_NO_GRAPH)
    cfg.wide_per_cu, cfg.narrow_per_cu = wide_per_cu, narrow_per_cu
    return capi.Plan(n, yield_capacity, cfg)


def evaluate(s: Surrogate, y, n: int = 1_000_000, seed: int = 5, reps: int = 5, log=None,
             fallback_per_cu: tuple[int, int] = (0, 0)) -> dict:
    """Surrogate + tree fallback against the plain FP64 tree on the mixed
    batch (95% inside the box, 5% outside): options/s of each (HIP events),
    the fallback fraction, and the error of every price against the tree's."""
    from . import datasets, devmem
    opts = datasets.box_mix(seed, n)
    dev = capi.to_device(opts, y)

    # plain tree at its best: the autotuned configuration, the plan's HIP
    # graph, timed with events on a caller stream
    tuned, _ = capi.autotune(dev)
    stream = C.c_void_p()
    devmem._ck(devmem.hip().hipStreamCreate(C.byref(stream)))
    plan = capi.Plan(n, y.N, tuned)
    plan.execute(dev, stream=stream.value)
    plan.sync()
    tree = dev["out"].to_host()
    e0, e1 = devmem.Event(), devmem.Event()
    e0.record(stream)
    for _ in range(reps):
        plan.execute(dev, stream=stream.value)
    e1.record(stream)
    e1.synchronize()
    plan.sync()
    tree_ms = e0.elapsed_ms(e1) / reps
    plan.close()

    fb_plan = fallback_plan(n, y.N, *fallback_per_cu)
    out = devmem.DeviceArray.empty((n,), np.float64)
    s.price(fb_plan, dev, out)  # warm-up
    runs = [s.price(fb_plan, dev, out) for _ in range(reps)]
    got = out.to_host()
    fb_plan.close()
    mean, spread, inbox = s.predict(dev)
    # the method without the network: the closed-form baseline alone (an
    # ensemble whose output layer is zero predicts exactly the baseline)
    P0 = s.get_params()
    P0[:, OFFSETS["W5"]:] = 0.0
    s0 = Surrogate(0, n, y.N, config(members=s.members))
    s0.set_params(P0)
    base, _, _ = s0.predict(dev)
    s0.close()
    devmem._ck(devmem.hip().hipStreamDestroy(stream))

    ms = float(np.mean([r["ms"] for r in runs]))
    err = np.abs(got - tree)
    surr_rows = inbox & (spread <= s.cfg.spread_threshold)
    err_net = np.abs(mean[inbox] - tree[inbox])
    rec = {
        "options": n,
        "tree_ms": tree_ms, "tree_options_per_s": n / (tree_ms * 1e-3),
        "surrogate_ms": ms, "surrogate_options_per_s": n / (ms * 1e-3),
        "speedup": tree_ms / ms,
        "tree_config": tuned.as_dict(),
        "inference_ms": float(np.mean([r["surrogate_ms"] for r in runs])),
        "fallback_tree_ms": float(np.mean([r["tree_ms"] for r in runs])),
        "fallback": int(runs[-1]["fallback"]), "out_of_box": int(runs[-1]["out_of_box"]),
        "fallback_frac": runs[-1]["fallback"] / n,
        "max_abs_err": float(err.max()), "mean_abs_err": float(err.mean()),
        "max_abs_err_surrogate_priced": float(err[surr_rows].max()) if surr_rows.any() else 0.0,
        "in_box_network_max_abs_err": float(err_net.max()) if err_net.size else 0.0,
        "in_box_network_mean_abs_err": float(err_net.mean()) if err_net.size else 0.0,
        "spread_threshold": float(s.cfg.spread_threshold),
        "fallback_per_cu": list(fallback_per_cu),
        "options_above_1e-3": int((err > 1e-3).sum()),
        "in_box_closed_form_max_abs_err": float(np.abs(base[inbox] - tree[inbox]).max()) if inbox.any() else 0.0,
        "in_box_closed_form_mean_abs_err": float(np.abs(base[inbox] - tree[inbox]).mean()) if inbox.any() else 0.0,
    }
    if log:
        log(json.dumps(rec))
    for k in list(dev):
        if hasattr(dev[k], "free"):
            dev[k].free()
    out.free()
    return rec
