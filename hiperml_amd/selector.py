"""Learned kernel selector: data generation, training and evaluation.

The pipeline (all timings and all training run on the GPU through libhml):
  1. random_shapes(): M, K log-uniform in [16, 65536], N log-uniform in
     [1, 512]; a third of the M, K and N draws rounded to a power of two
     and (M, K) another third to a multiple of 64, since real problems are
     mostly aligned; random transA/transB/beta==0, rejecting problems whose
     autotune working set exceeds the memory cap;
  2. autotune(): the exhaustive, validating autotuner (hmlSgemmAutotune)
     times every candidate variant of each shape into a CSV database;
  3. dataset(): per shape, the selector features, the mask of classes the
     autotuner measured as valid, and soft targets
     T_c ~ exp(-(t_c / t_best - 1) / tau) (tau = 0.05: a class 5% slower
     than the best gets weight 1/e);
  4. train(): the MLP (6 -> 3 x 64 ReLU -> classes) trained by the HIP
     kernels (hmlMlpTrain);
  5. evaluate(): on held-out shapes, the fraction of autotune-best
     throughput reached by the MLP's pick, by the built-in heuristic and by
     a nearest-bucket table built from the training shapes' winners.
"""
import csv
import ctypes
import math
import os

import numpy as np

from . import (GemmShape, Mlp, SgemmVariant, check, config_clear, config_load, lib,
               selector_install, sgemm_select, selector_class_of, selector_features,
               selector_mask, selector_variant)

TAU = 0.05


def autotune_bytes(M, N, K):
    # A, B, C0, C, reference C and the largest workspace (packed B + partials)
    return 4.0 * (M * K + K * N + 3 * M * N) + 4.0 * (K * max(N, 16) * 2 + 64 * M * N)


def _round_dim(rng, x, mult):
    """A third of the draws to the nearest power of two, a third (when
    mult > 1) to a multiple of `mult`, the rest unchanged."""
    u = rng.uniform()
    if u < 1.0 / 3.0:
        return 1 << max(0, int(round(math.log2(x))))
    if u < 2.0 / 3.0 and mult > 1 and x >= mult:
        return int(round(x / mult)) * mult
    return x


def random_shapes(n, seed, mem_cap=8e9, lo=16, hi=65536, nlo=1, nhi=512):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        M = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        K = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        N = int(round(math.exp(rng.uniform(math.log(nlo), math.log(nhi)))))
        M, K, N = _round_dim(rng, M, 64), _round_dim(rng, K, 64), _round_dim(rng, N, 1)
        if autotune_bytes(M, N, K) > mem_cap:
            continue
        ta, tb, b0 = (int(x) for x in rng.integers(0, 2, 3))
        out.append((M, N, K, ta, tb, b0))
    return out


# The shipped MI355X SGEMM table (hiperml_amd.tables writes it).
SHIPPED_TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "config", "hml_sgemm_mi355x.cfg")

# Row-major 8192 x N x 8192 problems, beta = 0: the skinny-SGEMM sweep.
NAMED_SHAPES = [(8192, n, 8192, 0, 0, 1) for n in (1, 2, 4, 8, 16, 32, 64, 128)]


def autotune(shapes, db_path, trials=3, flush=True, seed=1, progress=None):
    """Runs the exhaustive autotuner on every shape; returns the number of
    shapes that produced a winner."""
    ok = 0
    for i, (M, N, K, ta, tb, b0) in enumerate(shapes):
        s = GemmShape(M, N, K, 0, ta, tb, b0)
        best = SgemmVariant()
        bt, ht = ctypes.c_double(), ctypes.c_double()
        st = lib.hmlSgemmAutotune(ctypes.byref(s), trials, 1 if flush else 0, seed + i,
                                  ctypes.byref(best), ctypes.byref(bt), ctypes.byref(ht),
                                  db_path.encode())
        ok += st == 0
        if progress and (i + 1) % progress == 0:
            print("autotuned %d/%d shapes (last %dx%dx%d)" % (i + 1, len(shapes), M, N, K),
                  flush=True)
    return ok


def load_db(path):
    """{(M,N,K,transA,transB,beta0): {variant tuple: median ms}} (valid
    SGEMM rows only); `path` may list several databases separated by
    commas."""
    db = {}
    for one in str(path).split(","):
        with open(one) as f:
            for r in csv.DictReader(f):
                if r["kind"] != "0" or r["valid"] != "1":
                    continue
                key = (int(r["M"]), int(r["N"]), int(r["K"]), int(r["transA"]),
                       int(r["transB"]), int(r["beta0"]))
                v = (int(r["family"]), int(r["p1"]), int(r["p2"]), int(r["p3"]), int(r["p4"]))
                db.setdefault(key, {})[v] = float(r["median_ms"])
    return db


def shape_kw(key):
    M, N, K, ta, tb, b0 = key
    return dict(transA=bool(ta), transB=bool(tb), beta0=bool(b0))


def class_times(key, times):
    """Per selector class: measured median ms (inf where not measured).  A
    class whose variant for this shape equals a measured variant of another
    class (two geometries the split rule makes identical, a GEMV class that
    is the heuristic's own pick) takes that variant's time."""
    M, N, K = key[:3]
    kw = shape_kw(key)
    C = lib.hmlSelectorClassCount()
    t = np.full(C, np.inf)
    for v, ms in times.items():
        c = selector_class_of(M, N, K, v, **kw)
        if c >= 0:
            t[c] = min(t[c], ms)
    for c in np.nonzero(selector_mask(M, N, K, **kw) > 0)[0]:
        if not np.isfinite(t[c]):
            v = selector_variant(M, N, K, int(c), **kw)
            if v is not None and v.astuple() in times:
                t[c] = times[v.astuple()]
    return t


def dataset(db, tau=TAU):
    keys = sorted(db)
    X, T, Mk, TT = [], [], [], []
    for key in keys:
        M, N, K = key[:3]
        t = class_times(key, db[key])
        mask = (np.isfinite(t) & (selector_mask(M, N, K, **shape_kw(key)) > 0)).astype(np.float32)
        if mask.sum() == 0:
            continue
        tb = t[mask > 0].min()
        w = np.where(mask > 0, np.exp(-(np.where(mask > 0, t, tb) / tb - 1.0) / tau), 0.0)
        X.append(selector_features(M, N, K, **shape_kw(key)))
        T.append((w / w.sum()).astype(np.float32))
        Mk.append(mask)
        TT.append(t)
    return (np.array(keys), np.array(X, np.float32), np.array(T, np.float32),
            np.array(Mk, np.float32), np.array(TT))


def train(X, T, mask, epochs=300, batch=256, lr=2e-3, seed=7, hidden=64, layers=3):
    net = Mlp(X.shape[1], hidden, layers, T.shape[1], seed=seed)
    losses = net.train(X, T, mask, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, losses


def best_time(times):
    return min(times.values())


def evaluate(net, db_eval, db_train=None):
    """Fraction of autotune-best throughput (t_best / t_pick) per held-out
    shape for the MLP, the heuristic and (given training data) the
    nearest-bucket table of training winners."""
    keys, X, T, mask, tcls = dataset(db_eval)
    P = net.predict(X, mask) if net is not None else None
    res = {"mlp": [], "heuristic": [], "table": []}
    for i, key in enumerate(keys):
        key = tuple(int(v) for v in key)
        M, N, K, ta, tb, b0 = key
        times = db_eval[key]
        tbest = best_time(times)
        if P is not None:
            c = int(np.argmax(np.where(mask[i] > 0, P[i], -1.0)))
            res["mlp"].append(tbest / tcls[i][c])
        config_clear()
        h, _ = sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb), beta=0.0 if b0 else 1.0)
        res["heuristic"].append(tbest / times.get(h.astuple(), math.inf))
    if db_train is not None:
        heur = {}
        config_clear()
        for key in keys:
            key = tuple(int(v) for v in key)
            M, N, K, ta, tb, b0 = key
            heur[key] = sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb),
                                     beta=0.0 if b0 else 1.0)[0].astuple()
        install_table(db_train)
        for key in keys:
            key = tuple(int(v) for v in key)
            M, N, K, ta, tb, b0 = key
            times = db_eval[key]
            v, _ = sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb), beta=0.0 if b0 else 1.0)
            # a rule variant this shape's search never ran (a GEMV setting
            # of another bucket) counts as the heuristic's time
            t = times.get(v.astuple(), times.get(heur[key], math.inf))
            res["table"].append(best_time(times) / t)
        config_clear()
    return {k: np.array(v) for k, v in res.items() if v}


def install_table(db_train):
    """v2 config rules (one per shape bucket) from the training winners."""
    config_clear()
    for key, times in sorted(db_train.items()):
        M, N, K, ta, tb, b0 = key
        v = min(times, key=times.get)
        s = GemmShape(M, N, K, 0, ta, tb, b0)
        lib.hmlSgemmConfigAddRule(ctypes.byref(s), ctypes.byref(SgemmVariant(*v)))


def run(out_dir, n_train, n_eval, seed=2024, epochs=300, trials=3, mem_cap=8e9,
        train_db=None, eval_db=None, named=True):
    """End-to-end: autotune random training and held-out shapes (unless
    databases are given), train the MLP on the GPU, evaluate, save the model
    (hml_selector_mi355x.bin) and a JSON report into out_dir."""
    import json
    import os
    import time
    os.makedirs(out_dir, exist_ok=True)
    t0 = time.time()
    if train_db is None:
        train_db = os.path.join(out_dir, "autotune_train.csv")
        shapes = random_shapes(n_train, seed, mem_cap)
        autotune(shapes, train_db, trials=trials, seed=seed, progress=max(1, min(25, n_train // 10)))
        print("train autotune done in %.0f s" % (time.time() - t0), flush=True)
    if eval_db is None:
        eval_db = os.path.join(out_dir, "autotune_eval.csv")
        shapes = random_shapes(n_eval, seed + 1, mem_cap)
        autotune(shapes, eval_db, trials=trials, seed=seed + 100000,
                 progress=max(1, min(50, n_eval // 10)))
        print("eval autotune done in %.0f s" % (time.time() - t0), flush=True)
    named_db = None
    if named:
        named_db = os.path.join(out_dir, "autotune_named.csv")
        if os.path.exists(named_db):
            os.remove(named_db)
        autotune(NAMED_SHAPES, named_db, trials=trials, seed=seed + 200000)
    dbt, dbe = load_db(train_db), load_db(eval_db)
    keys, X, T, mask, _ = dataset(dbt)
    timings = sum(len(v) for v in dbt.values())
    print("training set: %d shapes, %d timings" % (len(keys), timings), flush=True)
    net, losses = train(X, T, mask, epochs=epochs, seed=seed)
    fr = evaluate(net, dbe, dbt)
    rep = {"train_shapes": len(keys), "train_timings": timings,
           "eval_shapes": len(dbe), "epochs": epochs,
           "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "fraction_of_autotune_best": summarize(fr),
           "named_shapes": named_report(net, load_db(named_db), dbt, SHIPPED_TABLE)
           if named_db else None,
           "seconds": round(time.time() - t0, 1)}
    net.save(os.path.join(out_dir, "hml_selector_mi355x.bin"))
    with open(os.path.join(out_dir, "selector_report.json"), "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps(rep), flush=True)
    return rep


def collect(out_dir, n, seed, name, trials=3, mem_cap=8e9, time_limit=None):
    """Autotunes `n` random shapes of stream `seed` into out_dir/name.csv
    (resuming: shapes already in that database cost nothing), stopping
    early after time_limit seconds."""
    import os
    import time
    os.makedirs(out_dir, exist_ok=True)
    db = os.path.join(out_dir, name + ".csv")
    shapes = random_shapes(n, seed, mem_cap)
    t0 = time.time()
    done = 0
    for i in range(0, len(shapes), 25):
        if time_limit and time.time() - t0 > time_limit:
            break
        done += autotune(shapes[i:i + 25], db, trials=trials, seed=seed * 1000003 + i)
        print("collected %d/%d shapes into %s (%.0f s)" % (min(i + 25, n), n, db,
                                                         time.time() - t0), flush=True)
    return done


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train and evaluate the learned kernel selector")
    ap.add_argument("--collect", metavar="NAME",
                    help="only autotune --train random shapes of stream --seed into "
                         "OUT/NAME.csv (resumable), then stop")
    ap.add_argument("--time-limit", type=float, help="with --collect: stop after this many seconds")
    ap.add_argument("--out", default="selector_out")
    ap.add_argument("--train", type=int, default=500, help="random training shapes")
    ap.add_argument("--eval", type=int, default=2000, help="random held-out shapes")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--trials", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--mem-cap", type=float, default=8e9)
    ap.add_argument("--train-db")
    ap.add_argument("--eval-db")
    ap.add_argument("--no-named", action="store_true",
                    help="skip the 8192 x N x 8192 shape report")
    a = ap.parse_args(argv)
    # the pipeline times the heuristic and the table itself: keep any
    # shipped selector out of this process's selections
    import os
    os.environ["HML_SELECTOR"] = "off"
    check(lib.hmlInit())
    if a.collect:
        collect(a.out, a.train, a.seed, a.collect, a.trials, a.mem_cap, a.time_limit)
        return
    run(a.out, a.train, a.eval, a.seed, a.epochs, a.trials, a.mem_cap, a.train_db, a.eval_db,
        not a.no_named)


def named_report(net, db_named, db_train, shipped_table=None):
    """Per named shape: the fraction of the autotune-best time reached by
    each selector; with `shipped_table` (a v2 .cfg file) also by the
    library's full selection order with that table loaded and `net`
    installed ("shipped": an exact-bucket rule, else the network)."""
    import os
    out = {}
    for key in sorted(db_named):
        fr = evaluate(net, {key: db_named[key]}, db_train)
        out["%dx%dx%d" % key[:3]] = {k: round(float(v[0]), 4) for k, v in fr.items() if len(v)}
    if shipped_table and os.path.exists(shipped_table) and net is not None:
        config_clear()
        config_load(shipped_table)
        selector_install(net)
        try:
            for key in sorted(db_named):
                M, N, K, ta, tb, b0 = key
                times = db_named[key]
                v, src = sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb),
                                      beta=0.0 if b0 else 1.0)
                t = times.get(v.astuple(), math.inf)
                out["%dx%dx%d" % key[:3]]["shipped"] = round(best_time(times) / t, 4)
        finally:
            selector_install(None)
            config_clear()
    return out


def summarize(fracs):
    out = {}
    for k, a in fracs.items():
        a = np.asarray(a, float)
        out[k] = {"mean": float(a.mean()), "geomean": float(np.exp(np.log(np.maximum(a, 1e-9)).mean())),
                  "min": float(a.min()), "within_5pct": float((a >= 0.95).mean()), "n": int(a.size)}
    return out


if __name__ == "__main__":
    main()
