"""Writes tests/golden/selector_heldout.npz: the held-out evaluation set of
the learned kernel selector, as measured on one MI355X by the exhaustive
autotuner (hmlSgemmAutotune over random shapes of a stream the training
shapes do not use; python -m hiperml_amd.selector --collect eval ...).

Stored per shape: (M, N, K, transA, transB, beta==0), the median time of
every selector class the autotuner measured (inf where it did not), the
best time of any measured variant, and the time of the built-in heuristic's
variant (selected with no table and no learned selector).  Tests score the
shipped selector's picks against these times without re-timing anything.

  python tests/golden/make_selector_fixture.py EVAL_DB.csv[,MORE.csv]
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))


def main(db_path, out=os.path.join(HERE, "selector_heldout.npz")):
    os.environ["HML_SELECTOR"] = "off"
    from hiperml_amd import config_clear, sgemm_select
    from hiperml_amd.selector import class_times, load_db
    config_clear()
    db = load_db(db_path)
    keys = sorted(db)
    shapes, cls_ms, best, heur = [], [], [], []
    for key in keys:
        M, N, K, ta, tb, b0 = key
        times = db[key]
        h, _ = sgemm_select(M, N, K, transA=bool(ta), transB=bool(tb), beta=0.0 if b0 else 1.0)
        th = times.get(h.astuple(), math.inf)
        if not math.isfinite(th):
            continue  # the heuristic's pick was not measured: not a usable row
        shapes.append(key)
        cls_ms.append(class_times(key, times).astype(np.float32))
        best.append(min(times.values()))
        heur.append(th)
    np.savez_compressed(out, shapes=np.array(shapes, np.int64),
                        class_ms=np.array(cls_ms, np.float32),
                        best_ms=np.array(best, np.float64),
                        heuristic_ms=np.array(heur, np.float64))
    print("wrote %s: %d shapes" % (out, len(shapes)))


if __name__ == "__main__":
    main(sys.argv[1])
