"""Writes published_results.json: the reference's published cuGAL and FUGAL
results for its benchmark experiment (benchmark/results/CUGAL_res.json,
benchmark/results/FUGAL_res2.json), one entry per graph and noise level:
the mean alignment time over the runs (Profile.time, the wall time of one
cugal() call, cugal/pred.py:243), the Sinkhorn phase time, the Sinkhorn
iterations of the first run and the mean node accuracy.  Run in the
container that holds the reference tree; the JSON is committed."""

import json
import os

REF = "/root/reference/benchmark/results"
HERE = os.path.dirname(os.path.abspath(__file__))


def summarise(path):
    data = json.load(open(path))
    exp = data["experiment"]
    out = {"experiment": exp, "commit": data.get("commit"), "entries": []}
    for gi, graph in enumerate(exp["graphs"]):
        for li, level in enumerate(exp["noise_levels"]):
            res = data["results"][gi][li][0]
            prof = res["profile"]
            its = [p["iteration_count"] for p in prof.get("sinkhorn_profiles", [])]
            out["entries"].append({
                "graph": graph["kind"], "noise": level, "time": prof["time"],
                "sinkhorn_time": prof.get("phase_times", {}).get("SINKHORN"),
                "sinkhorn_iterations": sum(its) if its else None, "solves": len(its) or None,
                "accuracy": res["accuracy"],
            })
    return out


if __name__ == "__main__":
    result = {"cugal": summarise(os.path.join(REF, "CUGAL_res.json")),
              "fugal_cpu": summarise(os.path.join(REF, "FUGAL_res2.json"))}
    with open(os.path.join(HERE, "published_results.json"), "w") as f:
        json.dump(result, f, indent=1)
    for e in result["cugal"]["entries"]:
        print(e["graph"], e["noise"]["source_noise"], round(e["time"], 2), round(e["accuracy"], 4),
              e["sinkhorn_iterations"], e["solves"])
