"""Learned style selector: an MLP (16 -> 64 -> 64 -> K, ReLU, FP32) over cheap
graph statistics that predicts, for each of an algorithm's K registered
styles, its slowdown against the fastest style; the pick is the style with
the smallest prediction.  Training and inference run on the GPU through the
hand-written HIP kernels of libindigo_hip.so (igx_mlp_*); the exhaustive
autotuner (autotune.sweep) is both the fallback and the yardstick of regret.

Targets: y_k = log(t_k / t_best), clipped at log(kClip); styles without a
measurement are masked out of the loss.  The library's default style is kept
unless another style's predicted slowdown is lower by more than a margin
chosen on held-back training graphs (fit_with_margin), so a wrong prediction
costs less where the default is already the best style.  Features are standardised with the
training set's mean and standard deviation, which are saved with the weights.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional

import numpy as np

from . import _native
from ._native import check

kClip = 100.0


def _lib():
    lib = _native.hip_lib()
    if not hasattr(lib, "_mlp_sigs"):
        vp = ctypes.c_void_p
        fp = ctypes.POINTER(ctypes.c_float)
        lib.igx_mlp_create.restype = ctypes.c_int
        lib.igx_mlp_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
        lib.igx_mlp_destroy.restype = ctypes.c_int
        lib.igx_mlp_destroy.argtypes = [vp]
        lib.igx_mlp_param_count.restype = ctypes.c_int
        lib.igx_mlp_param_count.argtypes = [vp]
        lib.igx_mlp_get_params.restype = ctypes.c_int
        lib.igx_mlp_get_params.argtypes = [vp, fp, ctypes.c_size_t]
        lib.igx_mlp_set_params.restype = ctypes.c_int
        lib.igx_mlp_set_params.argtypes = [vp, fp, ctypes.c_size_t]
        lib.igx_mlp_forward.restype = ctypes.c_int
        lib.igx_mlp_forward.argtypes = [vp, fp, ctypes.c_int, fp]
        lib.igx_mlp_gradients.restype = ctypes.c_int
        lib.igx_mlp_gradients.argtypes = [vp, fp, fp, fp, ctypes.c_int, ctypes.c_float, fp, fp]
        lib.igx_mlp_train.restype = ctypes.c_int
        lib.igx_mlp_train.argtypes = [vp, fp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_uint64, fp]
        lib._mlp_sigs = True
    return lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """Device MLP in -> 64 -> 64 -> out (libindigo_hip.so igx_mlp_*)."""

    HIDDEN = 64

    def __init__(self, n_in: int, n_out: int, seed: int = 1):
        self.n_in, self.n_out = n_in, n_out
        self.h = ctypes.c_void_p()
        check(_lib().igx_mlp_create(n_in, n_out, seed, ctypes.byref(self.h)), "igx_mlp_create")
        self.nparams = _lib().igx_mlp_param_count(self.h)

    def params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        check(_lib().igx_mlp_get_params(self.h, _fp(p), self.nparams), "igx_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(_lib().igx_mlp_set_params(self.h, _fp(p), p.size), "igx_mlp_set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], self.n_out), np.float32)
        check(_lib().igx_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)), "igx_mlp_forward")
        return y

    def gradients(self, x, t, mask, norm: Optional[float] = None):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        mask = np.ascontiguousarray(mask, np.float32)
        norm = float(mask.sum()) if norm is None else norm
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_float()
        check(_lib().igx_mlp_gradients(self.h, _fp(x), _fp(t), _fp(mask), x.shape[0], norm, _fp(g),
                                       ctypes.byref(loss)), "igx_mlp_gradients")
        return g, loss.value

    def train(self, x, t, mask, epochs: int, batch: int = 64, lr: float = 1e-3, seed: int = 1) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        mask = np.ascontiguousarray(mask, np.float32)
        loss = ctypes.c_float()
        check(_lib().igx_mlp_train(self.h, _fp(x), _fp(t), _fp(mask), x.shape[0], epochs, batch, lr, seed,
                                   ctypes.byref(loss)), "igx_mlp_train")
        return loss.value

    def free(self):
        if self.h:
            _lib().igx_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()


def unpack(p: np.ndarray, n_in: int, n_out: int):
    """Split a flat parameter vector into (W1, b1, W2, b2, W3, b3)."""
    H = Mlp.HIDDEN
    sizes = [n_in * H, H, H * H, H, H * n_out, n_out]
    parts = np.split(p.astype(np.float64), np.cumsum(sizes)[:-1])
    W1, b1, W2, b2, W3, b3 = parts
    return W1.reshape(n_in, H), b1, W2.reshape(H, H), b2, W3.reshape(H, n_out), b3


def targets_from_times(times: np.ndarray):
    """(y, mask): y = log(t / t_best) clipped at log(kClip).  +inf (a style
    too slow to run on that graph) gets the clipped maximum; NaN is masked."""
    t = np.asarray(times, np.float64)
    ok = np.isfinite(t) & (t > 0)
    slow = np.isposinf(t)
    best = np.where(ok, t, np.inf).min(axis=1, keepdims=True)
    y = np.where(ok, np.log(np.minimum(t / best, kClip)), 0.0)
    y = np.where(slow, np.log(kClip), y)
    return y.astype(np.float32), (ok | slow).astype(np.float32)


class Selector:
    def __init__(self, algo: str, styles: List[str], seed: int = 1, default: Optional[str] = None):
        self.algo = algo
        self.styles = list(styles)
        self.mean = None
        self.std = None
        self.net: Optional[Mlp] = None
        self.seed = seed
        # the library's default style and the predicted log-slowdown margin
        # another style must win by to be picked over it (fit_with_margin)
        self.default = default
        self.margin = 0.0

    def _norm(self, f: np.ndarray) -> np.ndarray:
        return ((np.asarray(f, np.float64) - self.mean) / self.std).astype(np.float32)

    def fit(self, features: np.ndarray, times: np.ndarray, epochs: int = 200, batch: int = 64,
            lr: float = 2e-3) -> float:
        f = np.asarray(features, np.float64)
        self.mean = f.mean(axis=0)
        self.std = np.where(f.std(axis=0) > 1e-12, f.std(axis=0), 1.0)
        y, m = targets_from_times(times)
        if self.net is None:
            self.net = Mlp(f.shape[1], len(self.styles), seed=self.seed)
        return self.net.train(self._norm(f), y, m, epochs=epochs, batch=batch, lr=lr, seed=self.seed)

    def predict_scores(self, features: np.ndarray) -> np.ndarray:
        return self.net.forward(self._norm(np.atleast_2d(features)))

    def pick(self, features: np.ndarray) -> List[int]:
        """Style index with the lowest predicted slowdown, among the styles
        the autotuner would run on that graph (feature 4 is log2(max degree +
        1); the sum of squared degrees follows from features 0, 2 and 3)."""
        from .autotune import sum_sq_degrees, too_slow
        f = np.atleast_2d(np.asarray(features, np.float64))
        scores = np.array(self.predict_scores(f), np.float64)
        for i in range(f.shape[0]):
            maxd = int(round(2.0 ** f[i, 4] - 1.0))
            sd2 = sum_sq_degrees(f[i])
            for k, st in enumerate(self.styles):
                if too_slow(self.algo, st, maxd, sd2):
                    scores[i, k] = np.inf
        picks = list(np.argmin(scores, axis=1))
        if self.default in self.styles and self.margin > 0.0:
            d = self.styles.index(self.default)
            for i, p in enumerate(picks):
                if np.isfinite(scores[i, d]) and scores[i, d] - scores[i, p] <= self.margin:
                    picks[i] = d
        return picks

    def fit_with_margin(self, features: np.ndarray, times: np.ndarray, default: str, epochs: int = 200,
                        batch: int = 64, lr: float = 2e-3, val_frac: float = 0.2,
                        margins=(0.0, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5), tol: float = 0.002) -> dict:
        """Fit, then choose the margin by which a style's predicted slowdown
        must beat the default's for it to be picked: the net is trained on a
        seeded (1 - val_frac) share of the graphs, the largest margin whose
        mean log slowdown against the best style on the rest is within `tol`
        of the lowest is kept, and the net is trained again on every graph."""
        f = np.asarray(features, np.float64)
        t = np.asarray(times, np.float64)
        self.default = default
        n = f.shape[0]
        order = np.random.default_rng(self.seed).permutation(n)
        nval = max(1, int(round(val_frac * n))) if n > 1 else 0
        val, tr = order[:nval], order[nval:]
        scores = {}
        if nval > 0 and len(tr) > 0:
            self.fit(f[tr], t[tr], epochs=epochs, batch=batch, lr=lr)
            tv = t[val]
            ok = np.isfinite(tv) & (tv > 0)
            best = np.where(ok, tv, np.inf).min(axis=1)
            for mg in margins:
                self.margin = mg
                pk = self.pick(f[val])
                chosen = tv[np.arange(len(val)), pk]
                chosen = np.where(np.isfinite(chosen) & (chosen > 0), chosen, best * kClip)
                scores[mg] = float(np.mean(np.log(chosen / best)))
            self.net.free()
            self.net = None
            # the most conservative margin whose validation regret is within
            # tol of the lowest: differences below that are timing noise
            low = min(scores.values())
            self.margin = max(mg for mg, v in scores.items() if v <= low + tol)
        loss = self.fit(f, t, epochs=epochs, batch=batch, lr=lr)
        return {"final_loss": loss, "margin": self.margin, "val_log_regret": scores}

    def evaluate(self, features: np.ndarray, times: np.ndarray, default_style: int) -> dict:
        t = np.asarray(times, np.float64)
        picks = self.pick(features)
        best = np.nanmin(np.where(np.isfinite(t), t, np.nan), axis=1)
        chosen = t[np.arange(t.shape[0]), picks]
        # a pick that was too slow to measure counts as the clipped slowdown
        chosen = np.where(np.isfinite(chosen) & (chosen > 0), chosen, best * kClip)
        regret = chosen / best - 1.0
        vs_default = t[:, default_style] / chosen
        return {"mean_regret": float(np.mean(regret)), "median_regret": float(np.median(regret)),
                "max_regret": float(np.max(regret)), "frac_within_10pct": float(np.mean(regret <= 0.10)),
                "geomean_speedup_vs_default": float(np.exp(np.mean(np.log(vs_default)))),
                "frac_faster_than_default": float(np.mean(vs_default > 1.0)),
                "picks": [self.styles[p] for p in picks]}

    def save(self, path: str) -> None:
        np.savez_compressed(path, algo=self.algo, styles=np.array(self.styles), mean=self.mean, std=self.std,
                            params=self.net.params(), n_in=self.net.n_in, default=self.default or "",
                            margin=self.margin)

    @classmethod
    def load(cls, path: str) -> "Selector":
        d = np.load(path)
        s = cls(str(d["algo"]), [str(x) for x in d["styles"]])
        s.mean, s.std = d["mean"], d["std"]
        s.net = Mlp(int(d["n_in"]), len(s.styles))
        s.net.set_params(d["params"])
        if "default" in d.files and str(d["default"]):
            s.default = str(d["default"])
            s.margin = float(d["margin"])
        return s


MODEL_DIR = os.path.join(os.path.dirname(__file__), "selector_models")
_loaded = {}


def pick_style(graph, algo: str, model_dir: Optional[str] = None) -> str:
    """Style name the trained selector picks for `graph` (a Graph) and `algo`
    ("bfs", "sssp", "cc", "mis", "pr" or "tc"), from the models shipped in
    selector_models/ unless model_dir names another directory.  Needs the GPU
    library (the MLP forward pass runs on the device)."""
    sel = _model(algo, model_dir)
    return sel.styles[sel.pick(graph.features())[0]]


def _model(algo: str, model_dir: Optional[str] = None) -> "Selector":
    d = model_dir or MODEL_DIR
    key = (d, algo)
    if key not in _loaded:
        _loaded[key] = Selector.load(os.path.join(d, f"selector_{algo}.npz"))
    return _loaded[key]


def pick_style_device(dg, algo: str, model_dir: Optional[str] = None, seed: int = 1):
    """Like pick_style, for a graph already on the device (a DeviceGraph):
    the features are computed there (igx_graph_features_device), so the CSR
    is not copied back.  Returns (style name, feature seconds)."""
    from . import device
    f, secs = device.features_device(dg, seed)
    sel = _model(algo, model_dir)
    return sel.styles[sel.pick(f)[0]], secs
