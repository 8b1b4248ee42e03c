"""Collect style timings, train the per-algorithm selectors on the GPU, and
evaluate them on held-out graphs.

  python -m indigo2suite_amd.train_selector collect --count N --seed-base S \
        [--log-n LO HI] [--deg LO HI] --out FILE.npz
  python -m indigo2suite_amd.train_selector retime --data FILE.npz --out FILE2.npz --algo A --styles S [S ...] \
        --seed-base S [--log-n LO HI] [--deg LO HI]
  python -m indigo2suite_amd.train_selector fit --train FILE.npz [FILE.npz ...] --out-dir DIR
  python -m indigo2suite_amd.train_selector eval --test FILE.npz [FILE.npz ...] --model-dir DIR [--json OUT]

`collect` runs the exhaustive autotuner (every registered style of every
algorithm, median of three runs) on seeded graphs from four families;
`retime` times some styles again on a dataset's graphs (after their kernels
changed) and replaces their columns; `fit` trains one MLP per algorithm with the HIP kernels; `eval` reports, per
algorithm, the regret of the learned pick against the exhaustive best and its
speedup over the fixed default style (the library's entry style).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import autotune, device
from .selector import Selector

ALGOS = ("bfs", "sssp", "cc", "mis", "pr", "tc")


def cmd_collect(a):
    t0 = time.time()
    data = autotune.collect(a.count, ALGOS, seed_base=a.seed_base, log_n_range=tuple(a.log_n),
                            deg_range=tuple(a.deg), repeats=a.repeats,
                            log=lambda s: print(s, flush=True),
                            checkpoint=lambda d: autotune.save_dataset(a.out, d),
                            checkpoint_every=a.checkpoint_every, max_seconds=a.max_seconds)
    autotune.save_dataset(a.out, data)
    print(f"wrote {a.out}: {len(data['features'])} graphs in {time.time() - t0:.0f}s", flush=True)


def cmd_retime(a):
    t0 = time.time()
    n = autotune.retime(a.data, a.out, a.algo, a.styles, a.seed_base, tuple(a.log_n), tuple(a.deg),
                        repeats=a.repeats, log=lambda s: print(s, flush=True))
    print(f"wrote {a.out}: {a.algo} {a.styles} retimed on {n} graphs in {time.time() - t0:.0f}s", flush=True)


def load_many(paths):
    """Concatenate datasets collected with the same style registry."""
    parts = [autotune.load_dataset(p) for p in paths]
    styles = parts[0]["styles"]
    for d in parts[1:]:
        assert d["styles"] == styles, "datasets from different style registries"
    return {"features": np.concatenate([d["features"] for d in parts]),
            "times": {a: np.concatenate([d["times"][a] for d in parts]) for a in styles}, "styles": styles}


def cmd_fit(a):
    data = load_many(a.train)
    os.makedirs(a.out_dir, exist_ok=True)
    summary = {}
    device.init()
    for algo in ALGOS:
        sel = Selector(algo, data["styles"][algo], seed=a.seed)
        t = time.time()
        default = device.styles(algo)[device.default_entry_style(algo, 0)]
        res = sel.fit_with_margin(data["features"], data["times"][algo], default, epochs=a.epochs,
                                  batch=a.batch, lr=a.lr)
        sel.save(os.path.join(a.out_dir, f"selector_{algo}.npz"))
        summary[algo] = {**res, "default": default, "train_seconds": round(time.time() - t, 2),
                         "graphs": int(data["features"].shape[0])}
        print(algo, summary[algo], flush=True)
        sel.net.free()
    print(json.dumps(summary), flush=True)


def cmd_eval(a):
    data = load_many(a.test)
    device.init()
    out = {}
    for algo in ALGOS:
        sel = Selector.load(os.path.join(a.model_dir, f"selector_{algo}.npz"))
        assert sel.styles == data["styles"][algo], "style registry changed since training"
        default = device.lib().igx_get_entry_style(device.ALGOS[algo], 0)
        ev = sel.evaluate(data["features"], data["times"][algo], default_style=default)
        t = np.asarray(data["times"][algo])
        best = np.argmin(t, axis=1)
        tb = t[np.arange(len(best)), best]
        ev["default_style"] = sel.styles[default]
        ev["default_regret_mean"] = float(np.mean(t[:, default] / tb - 1.0))
        ev["best_speedup_vs_default"] = float(np.exp(np.mean(np.log(t[:, default] / tb))))
        ev["margin"] = sel.margin
        ev["best_styles"] = [sel.styles[b] for b in best]
        out[algo] = ev
        print(algo, {k: v for k, v in ev.items() if not isinstance(v, list)}, flush=True)
        sel.net.free()
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
    if a.markdown:
        with open(a.markdown, "w") as f:
            f.write(markdown_table(out))


def markdown_table(ev: dict) -> str:
    """One row per algorithm: the learned pick's regret against the tuned best,
    the default's, and speedups over the default (geometric means)."""
    lines = ["| algorithm | default style | margin | mean regret, learned | max | within 10% of best | "
             "mean regret, default | speedup vs default, learned | speedup vs default, tuned best | "
             "faster than default on |",
             "|---|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for algo, e in ev.items():
        lines.append(f"| {algo} | `{e['default_style']}` | {e['margin']:.2f} | {100 * e['mean_regret']:.1f}% | "
                     f"{100 * e['max_regret']:.0f}% | {100 * e['frac_within_10pct']:.0f}% | "
                     f"{100 * e['default_regret_mean']:.1f}% | {e['geomean_speedup_vs_default']:.3f} | "
                     f"{e['best_speedup_vs_default']:.3f} | {100 * e['frac_faster_than_default']:.0f}% |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--count", type=int, required=True)
    c.add_argument("--seed-base", type=int, default=100000)
    c.add_argument("--log-n", type=float, nargs=2, default=(12, 21))
    c.add_argument("--deg", type=float, nargs=2, default=(4, 32))
    c.add_argument("--repeats", type=int, default=3)
    c.add_argument("--out", required=True)
    c.add_argument("--checkpoint-every", type=int, default=25, help="save the partial dataset every N graphs")
    c.add_argument("--max-seconds", type=float, default=0.0, help="stop after the graph that passes this wall time")
    r = sub.add_parser("retime", help="time some styles again on a dataset's graphs, replacing their columns")
    r.add_argument("--data", required=True)
    r.add_argument("--out", required=True)
    r.add_argument("--algo", required=True, choices=ALGOS)
    r.add_argument("--styles", required=True, nargs="+")
    r.add_argument("--seed-base", type=int, required=True)
    r.add_argument("--log-n", type=float, nargs=2, default=(12, 21))
    r.add_argument("--deg", type=float, nargs=2, default=(4, 32))
    r.add_argument("--repeats", type=int, default=3)
    f = sub.add_parser("fit")
    f.add_argument("--train", required=True, nargs="+")
    f.add_argument("--out-dir", required=True)
    f.add_argument("--epochs", type=int, default=300)
    f.add_argument("--batch", type=int, default=64)
    f.add_argument("--lr", type=float, default=2e-3)
    f.add_argument("--seed", type=int, default=1)
    e = sub.add_parser("eval")
    e.add_argument("--test", required=True, nargs="+")
    e.add_argument("--model-dir", required=True)
    e.add_argument("--json")
    e.add_argument("--markdown", help="write the per-algorithm table as markdown")
    a = ap.parse_args(argv)
    {"collect": cmd_collect, "retime": cmd_retime, "fit": cmd_fit, "eval": cmd_eval}[a.cmd](a)


if __name__ == "__main__":
    main(sys.argv[1:])
