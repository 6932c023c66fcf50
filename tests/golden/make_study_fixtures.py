"""Writes tests/golden/study_cases.json from the reference's dense study
configurations and result tables (data only: case definitions, solver
limits and, per instance, the AMFLS row's seeds, reference objective and
outcome).  Reads /root/reference; run in the build container:

    python tests/golden/make_study_fixtures.py
"""
import csv
import json
from pathlib import Path

REF = Path("/root/reference/experiments")
SUITES = ["large_dense_ordinary_v2", "large_dense_ridge_v1", "large_dense_geometry_rank_v1",
          "large_dense_spectral_coverage_v1"]
KEEP = ["tolerance", "failure_probability", "algorithm_seed_offset", "maximum_iterations", "maximum_depth",
        "maximum_auxiliary_width", "maximum_epochs", "maximum_basis_size"]
ROW = ["instance_seed", "algorithm_seed", "reference_objective", "external_success", "status", "depth",
       "relative_solution_error", "relative_energy_error_external", "backward_error_external"]


def main():
    out = {}
    for suite in SUITES:
        cfg = json.loads((REF / "configs" / f"{suite}.json").read_text())
        rows = [r for r in csv.DictReader((REF / "results" / "formal_100_v2" / f"{suite}.csv").open())
                if r["method"] == "amfls"]
        cases = []
        for case in cfg["cases"]:
            mine = [r for r in rows if r["case"] == case["name"]]
            mine.sort(key=lambda r: int(r["instance_seed"]))
            cases.append({"case": case, "instances": [{k: r[k] for k in ROW} for r in mine[:3]]})
        out[suite] = {"options": {k: cfg[k] for k in KEEP}, "cases": cases}
    path = Path(__file__).resolve().parent / "study_cases.json"
    path.write_text(json.dumps(out, indent=1))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
