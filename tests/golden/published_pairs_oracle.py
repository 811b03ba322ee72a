"""Writes published_pairs_oracle.json: the FP64 oracle's alignment of the
reference's published experiment pairs (benchmark/results/CUGAL_res.json:
the shipped graphs, seed 283841238, the pairs drawn in the order of
benchmark/experiment.py:440-460) at source noise 0.1 with the published
cuGAL Config (LOG, reg 0.3, <= 500 iterations, thr 1e-3, mu 2, 15 x 10 FW,
sparse adjacency, SCIPY).  Per run: node accuracy and every solve's
Sinkhorn iteration count.  With --quirk the oracle also applies the CUDA
branch's in-place marginal check of cugal/sinkhorn.py:67-74 (called at
:112-113, with the v-update reading the copy of K taken at :99-100).

    python tests/golden/published_pairs_oracle.py IN_ARENAS 0,1,2,3 [--quirk]

Torch's CPU reductions split by thread count, so the count is fixed (4) and
recorded."""

import argparse
import copy
import fcntl
import json
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cugal_amd.experiment import DATA_FILES, GraphKind, NoiseLevel, generate_graph, load_edge_list  # noqa: E402
from oracle import fugal as oracle  # noqa: E402

THREADS = 4
OUT = os.path.join(HERE, "published_pairs_oracle.json")


def published_pairs():
    exp = json.load(open(os.path.join(HERE, "published_results.json")))["cugal"]["experiment"]
    levels = [NoiseLevel(**lv) for lv in exp["noise_levels"]]
    gen = np.random.default_rng(seed=exp["seed"])
    pairs = {}
    for g in exp["graphs"]:
        kind = GraphKind[g["kind"]]
        graph = load_edge_list(os.path.join(HERE, DATA_FILES[kind]))
        drawn = [[generate_graph(graph, None, gen, lv) for _ in range(exp["num_runs"])] for lv in levels]
        pairs[kind.value] = drawn[0]
    return pairs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("graph")
    ap.add_argument("runs")
    ap.add_argument("--quirk", action="store_true")
    args = ap.parse_args()
    torch.set_num_threads(THREADS)
    pairs = published_pairs()
    cfg = oracle.OracleConfig(sinkhorn_method="LOG", sinkhorn_iterations=500, hungarian_method="SCIPY",
                              sinkhorn_regularization=0.3, sinkhorn_threshold=1e-3, mu=2.0, iter_count=15,
                              frank_wolfe_iter_count=10, use_sparse_adjacency=True,
                              quirk_log_marginal="gpu" if args.quirk else False)
    key = args.graph + (" quirk" if args.quirk else "")
    for r in [int(x) for x in args.runs.split(",")]:
        s, t, mapping, _ = pairs[args.graph][r]
        t0 = time.time()
        res = oracle.align(copy.deepcopy(s), copy.deepcopy(t), cfg)
        acc = float(np.mean(np.array([m for _, m in res.mapping]) == np.asarray(mapping)))
        with open(OUT + ".lock", "w") as lock:  # several graphs may run at once
            fcntl.flock(lock, fcntl.LOCK_EX)
            data = json.load(open(OUT)) if os.path.exists(OUT) else {"threads": THREADS, "graphs": {}}
            data["graphs"].setdefault(key, {})[str(r)] = {"accuracy": acc,
                                                          "sinkhorn_iterations": list(res.sinkhorn_iterations)}
            with open(OUT, "w") as f:
                json.dump(data, f, indent=0, sort_keys=True)
        print(f"{key} run {r}: accuracy {acc:.4f}, {sum(res.sinkhorn_iterations)} iterations "
              f"({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
