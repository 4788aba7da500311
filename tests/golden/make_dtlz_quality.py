"""Writes dtlz_quality_reference.csv: the reference's published final IGD of
every CUDA-MOEA DTLZ quality run (NSGA-III and RVEA; DTLZ1-7 and ConvexDTLZ2;
M = 3, D = 7 for DTLZ1 and 12 otherwise, N = 1024 (RVEA 990), 500 generations,
seeds 0-29; release 0.1.0 of the reference), copied from the reference's
result table tests/benchmark/DTLZ/results/data/quality_runs.csv.  Data only: run where the
reference checkout is present, at /root/reference or the path given."""
import csv
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
SRC = REF / "tests/benchmark/DTLZ/results/data/quality_runs.csv"
OUT = Path(__file__).resolve().parent / "dtlz_quality_reference.csv"
FIELDS = ["problem", "algorithm", "seed", "framework_version", "dimension", "generations", "actual_population",
          "active_population",
          "reference_direction_count", "final_igd", "igd_20000", "igd_50000", "pareto_reference_points"]

with SRC.open(newline="") as f:
    rows = [r for r in csv.DictReader(f) if r["framework"] == "cuda_moea" and r["run_status"] == "success"]
rows.sort(key=lambda r: (r["problem"], r["algorithm"], int(r["seed"])))
with OUT.open("w", newline="") as f:
    w = csv.DictWriter(f, fieldnames=FIELDS)
    w.writeheader()
    for r in rows:
        w.writerow({k: r[k] for k in FIELDS})
print(f"{len(rows)} rows -> {OUT}")
