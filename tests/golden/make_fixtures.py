"""Extract the reference-held data files that pin the oracle into compact
fixtures under tests/golden/.  Run once in the build container (it reads the
reference repository); the fixtures are committed so tests never need it.

Sources (data files held by the reference, not code):
  web/public/data/benchmarks/series6-accuracy.json  (blaze_f64 / mpb band
      frequencies omega/2pi; square lattice, eps_bg=1, rods eps=8.9 r=0.2 at
      (0.5,0.5); 32x32; 20 bands computed, lowest 10 kept; 15 k per leg; tol 1e-7)
  web/public/data/epsilon-grids/epsilon-res{16,32,48,64}.json  (smoothed eps
      exported by the reference CLI: square, eps_bg=13, hole eps=1 r=0.3 at
      (0.5,0.5), analytic smoothing)
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    s6 = json.load(open(os.path.join(REF, "web/public/data/benchmarks/series6-accuracy.json")))
    fx = {"parameters": s6["parameters"]}
    for pol in ("TM", "TE"):
        fx[pol] = {
            "blaze_f64": [k["frequencies"] for k in s6[pol]["blaze_f64"]],
            "blaze_f32": [k["frequencies"] for k in s6[pol]["blaze_f32"]],
            "mpb": [k["frequencies"] for k in s6[pol]["mpb"]],
            "k_distance": [k["k_distance"] for k in s6[pol]["blaze_f64"]],
        }
    with open(os.path.join(OUT, "series6_accuracy.json"), "w") as f:
        json.dump(fx, f)
    grids = {}
    for res in (16, 32, 48, 64):
        d = json.load(open(os.path.join(REF, f"web/public/data/epsilon-grids/epsilon-res{res}.json")))
        grids[str(res)] = {k: d[k] for k in ("nx", "ny", "epsilon_background", "epsilon_hole", "radius", "grid")}
    with open(os.path.join(OUT, "epsilon_grids.json"), "w") as f:
        json.dump(grids, f)


if __name__ == "__main__":
    main()
