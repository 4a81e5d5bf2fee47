"""Extract golden vectors (data files only) from the reference's recorded runs.

Run in a container that has the reference checkout; the outputs are committed
so the tests never need it.  Sources (reference paths):
  docs/verification_stage1_20261106/corrected32_t12/{project.json,history.csv}
  docs/verification_farfield_20261107/airfoil/characteristic_down3/{project.json,history.csv}
  docs/verification_stage1_20261106/profile_threads1.json (last_sample after 272 steps)
  tests/fixtures/legacy051_project.json (fingerprint fixture)
  docs/verification_040/{startup_project.json,startup_history.csv} (two-level startup, every 4 coarse steps)
  docs/verification_multielement_050/{fowler,double,slotted}/{project.json,history.csv,residuals.csv,summary.json}
      (two- and three-element cases, every coarse step; summary diagnostics: budget and guards)
  docs/verification_user_tests_20261108/raw/final_smoke/study/cases/N16/{project.json,history.csv}
      (exterior farfield, 136x51x8, every step from step 1)
  docs/verification_exterior_library_20261108/raw/cold_offline_final_v2/runtime/parent/{project.json,history.csv}
      (exterior farfield, 88x51x8, every 4 steps)
  docs/verification_exterior_library_20261108/raw/matched_airfoils/airfoil_down3/{project.json,history.csv}
      (exterior farfield, 176x101x16, every 8 steps)
"""
import csv
import json
import shutil
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent

CASES = {
    "n32_legacy_down3": ("docs/verification_stage1_20261106/corrected32_t12", 240),
    "n32_characteristic_down3": ("docs/verification_farfield_20261107/airfoil/characteristic_down3", 400),
    "exterior_n16_smoke": ("docs/verification_user_tests_20261108/raw/final_smoke/study/cases/N16", 48),
    "exterior_n16_short": ("docs/verification_exterior_library_20261108/raw/cold_offline_final_v2/runtime/parent", 32),
    "exterior_n32_down3": ("docs/verification_exterior_library_20261108/raw/matched_airfoils/airfoil_down3", 400),
}

for name, (rel, last_step) in CASES.items():
    dst = OUT / name
    dst.mkdir(exist_ok=True)
    shutil.copyfile(REF / rel / "project.json", dst / "project.json")
    with open(REF / rel / "history.csv", newline="") as src, open(dst / "history_head.csv", "w", newline="") as out:
        reader = csv.DictReader(src)
        writer = csv.DictWriter(out, reader.fieldnames)
        writer.writeheader()
        for row in reader:
            if int(row["step"]) > last_step:
                break
            writer.writerow(row)

REFINED = {
    "refined_startup_n32": ("docs/verification_040", "startup_project.json", "startup_history.csv", 40),
    "refined_fowler_n32": ("docs/verification_multielement_050/fowler", "project.json", "history.csv", 32),
    "refined_double_n32": ("docs/verification_multielement_050/double", "project.json", "history.csv", 32),
    "refined_slotted_n32": ("docs/verification_multielement_050/slotted", "project.json", "history.csv", 800),
}
for name, (rel, proj, hist, last_step) in REFINED.items():
    dst = OUT / name
    dst.mkdir(exist_ok=True)
    shutil.copyfile(REF / rel / proj, dst / "project.json")
    with open(REF / rel / hist, newline="") as src, open(dst / "history_head.csv", "w", newline="") as out:
        reader = csv.DictReader(src)
        writer = csv.DictWriter(out, reader.fieldnames)
        writer.writeheader()
        for row in reader:
            if int(row["step"]) > last_step:
                break
            writer.writerow(row)

# The two-level runs' conservation budget and residuals: summary.json "diagnostics" (guards,
# cumulative wall mass exchange, last_budget) and residuals.csv (l2/linf/balance columns).
for name in ("refined_fowler_n32", "refined_double_n32", "refined_slotted_n32"):
    rel = REFINED[name][0]
    summary = json.loads((REF / rel / "summary.json").read_text())
    (OUT / name / "summary_diagnostics.json").write_text(json.dumps(
        {"steps": summary["steps"], "diagnostics": summary["diagnostics"], "mean": summary["mean"]}, indent=1) + "\n")
    shutil.copyfile(REF / rel / "residuals.csv", OUT / name / "residuals.csv")

profile = json.loads((REF / "docs/verification_stage1_20261106/profile_threads1.json").read_text())
(OUT / "profile_n32_naca0012_81.json").write_text(json.dumps({
    "case": {"naca0012_n_side": 81, "chord": 1.0, "moment_reference": [0.25, 0.0],
             "flow": {"reynolds": 30000, "alpha_deg": 4, "physical_chord_m": 0.15},
             "grid": {"cells_per_chord": 32, "upstream_chords": 1.5, "downstream_chords": 3,
                      "transverse_chords": 1.5, "span_chords": 0.5, "memory_limit_gib": 2},
             "farfield": "legacy", "steps": 272, "sample_every": 8},
    "shape": profile["shape"], "last_sample": profile["last_sample"]}, indent=2) + "\n")
shutil.copyfile(REF / "tests/fixtures/legacy051_project.json", OUT / "legacy051_project.json")
