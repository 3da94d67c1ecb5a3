"""Autotuner of the MI355X plugin's launch shapes and queue capacity.

Sweeps the run-time tuning keys of libprender_gpu.so (prgpu_tune_set) on one GPU and writes the
fastest setting to the tuning file the plugin reads when it first enumerates devices
(prgpu_tuning.txt beside the library, or --out):

  trace_blocks_per_cu   persistent traversal blocks per CU (the ray-fetch wavefront width)
  shade_blocks_per_cu   blocks per CU of the surface, emission and material kernels
  path_budget           paths in flight per batch: the capacity of every ray and material queue
  graph_rounds          bounce rounds replayed per captured HIP graph (0: eager launches)
  material_sort         sort the material queues by material and medium each bounce (0 / 1)
  ray_sort              sort the next bounce's rays by direction octant and origin cell (0 / 1)

The objective is the wall time of render calls on a benchmark configuration (default: the
1920x1080 terrain scene, configuration 2, rendered into the
device-resident film; 16 spp per call, four times the bench's step, so that a call spans several
batches at the smaller queue capacities). Coordinate descent: each key in turn is set to every candidate with the
others held at their current best, for --passes passes. Every candidate renders the same sample
ranges, and the film of each candidate is checked against the first one's (a launch shape never
changes a sample's path, nor does the order the material stage takes its queues in; a smaller
queue capacity splits a call into more batches, which only
regroups the FP64 film additions: relative 1e-9), so a setting that rendered differently is never
chosen.

Usage (on a GPU): python ar-renderer_amd/autotune.py [--config 2] [--calls 3] [--out FILE]
"""
import argparse
import datetime
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import prender_amd as pa  # noqa: E402

CANDIDATES = {
    "trace_blocks_per_cu": [4, 5, 6, 7, 8],
    "shade_blocks_per_cu": [4, 8, 12, 16, 24, 32],
    "path_budget": [1 << 22, 1 << 23, 1 << 24, 1 << 25],
    "graph_rounds": [0, 2, 4],
    "material_sort": [0, 1],
    "ray_sort": [0, 1],
}
# (config, width, height, max depth, rr start, spp per call). Configuration 2 is swept at 16 spp per
# call (33 M samples): a call larger than the smaller queue capacities, so that the capacity's cost
# (a batch's late, nearly empty bounces, paid once per batch) is part of the objective.
WORKLOADS = {1: (1, 256, 256, 8, 3, 16), 2: (2, 1920, 1080, 16, 1, 16), 3: (3, 1280, 720, 256, 1, 4),
             4: (4, 1920, 1080, 32, 1, 4), 5: (5, 3840, 2160, 16, 1, 1)}


def measure(plugin, scene, w, h, depth, rr_start, spp, calls, first):
    """Seconds per call over `calls` calls after one warm-up call, and the film of the timed calls."""
    plugin.film_reset(w, h)
    prm = lambda f: pa.render_params(spp=spp, first_sample=f, seed=1, max_depth=depth, rr_start=rr_start)  # noqa: E731
    plugin.render_resident(scene, prm(first))
    plugin.film_reset(w, h)
    t0 = time.perf_counter()
    for k in range(calls):
        plugin.render_resident(scene, prm(first + spp * (k + 1)))
    el = (time.perf_counter() - t0) / calls
    film, _ = plugin.film_read(w, h, variance=False)
    return el, film


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--config", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--calls", type=int, default=3)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(HERE, "prgpu_tuning.txt"))
    args = ap.parse_args()

    host, plugin = pa.Host(), pa.Plugin()
    devs = plugin.devices()
    if not devs:
        raise SystemExit("autotune: no HIP device visible")
    cfg, w, h, depth, rr_start, spp = WORKLOADS[args.config]
    scene = host.generate(cfg, w, h)
    best = {k: plugin.tune_get(k) for k in CANDIDATES}
    log = []
    ref_film = None
    t_best = None
    for p in range(args.passes):
        for key, values in CANDIDATES.items():
            results = {}
            for v in values:
                for k2, v2 in best.items():
                    plugin.tune_set(k2, v2)
                plugin.tune_set(key, v)
                el, film = measure(plugin, scene, w, h, depth, rr_start, spp, args.calls, 1000)
                if ref_film is None:
                    ref_film = film
                same = bool(np.allclose(film, ref_film, rtol=1e-9, atol=0))
                results[v] = el
                log.append({"pass": p, "key": key, "value": v, "ms_per_call": round(el * 1e3, 3), "same_film": same})
                print(json.dumps(log[-1]), flush=True)
                if not same:
                    results[v] = float("inf")
            v_best = min(results, key=results.get)
            best[key] = v_best
            t_best = results[v_best]
    for k2, v2 in best.items():
        plugin.tune_set(k2, v2)
    lines = [f"# prender_gpu tuning, written by ar-renderer_amd/autotune.py on {datetime.date.today().isoformat()}",
             f"# device: {devs[0].name.decode()}",
             f"# workload: configuration {cfg} at {w}x{h}, max depth {depth}, {spp} spp per call; "
             f"best {t_best * 1e3:.3f} ms per call",
             f"# candidates: {json.dumps(CANDIDATES)}"]
    lines += [f"{k} {int(v) if float(v).is_integer() else v}" for k, v in best.items()]
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(json.dumps({"tuning": best, "ms_per_call": round(t_best * 1e3, 3), "file": args.out}), flush=True)


if __name__ == "__main__":
    main()
