"""Writes tests/golden/reference_goldens.json: the golden (iterations, RMSE)
values that the reference's own tests and benchmark tables hold.

Run in the build container only (it reads the reference's files as text; the
reference is never imported or executed):

    python tests/golden/make_goldens.py /root/reference

Sources:
  * examples/benchmark.py:26-632 -- dictionaries reference_v, reference_vskip,
    reference_f, reference_fskip, reference_w, reference_wskip mapping
    (problem, resolution) -> (iterations, rmse) for 2D square grids.
  * tests/test_solver_2d.py and tests/test_solver_3d.py -- asserted iteration
    counts and RMSE values per test function (transcribed below with the
    test function and line each value comes from).
"""
import json
import math
import os
import re
import sys

TABLES = ["reference_v", "reference_vskip", "reference_f", "reference_fskip",
          "reference_w", "reference_wskip"]

# tests/test_solver_2d.py: (problem, mx, my) -> (iterations, rmse); line of the assert
TESTS_2D = [
    (1, 128, 128, 13, 1.010315424011278e-06, 170),
    (1, 64, 64, 13, 4.010172727197344e-06, 218),
    (1, 64, 128, 27, 4.025383882544280e-06, 235),
    (2, 128, 128, 14, 2.490588070275737e-05, 278),
    (2, 64, 64, 14, 9.886612163935591e-05, 295),
    (2, 64, 128, 28, 6.202644125853758e-05, 312),
    (3, 128, 128, 41, 2.197736631958706e-05, 357),
    (3, 64, 64, 35, 8.725944838558505e-05, 374),
    (3, 64, 128, 77, 5.910207160483823e-05, 391),
    (4, 128, 128, 468, 2.179081815661214e-05, 437),
    (4, 64, 64, 228, 8.596967777404521e-05, 454),
    (4, 64, 128, 796, 5.979198934671856e-05, 471),
    (5, 128, 128, 208, 2.483558318071642e-05, 516),
    (5, 64, 64, 205, 9.858702741615382e-05, 533),
    (5, 64, 128, 107, 9.823336470021998e-05, 550),
    (7, 128, 128, 466, 2.291672360185362e-05, 599),
    (7, 64, 64, 226, 8.920084594712214e-05, 616),
    (7, 64, 128, 794, 6.078926956740372e-05, 633),
    (9, 128, 128, 462, 2.719754537531388e-05, 682),
    (9, 64, 64, 224, 1.073853922527805e-04, 699),
    (9, 64, 128, 785, 7.061117869397386e-05, 716),
    (10, 128, 128, 41, 2.317202758890556e-05, 765),
    (10, 64, 64, 34, 9.205560538691322e-05, 782),
    (10, 64, 128, 33, 6.674092567109719e-05, 799),
    (11, 128, 128, 56, 3.881907960699822e-05, 862),
    (11, 64, 64, 53, 1.541913423109781e-04, 879),
    (11, 64, 128, 155, 1.008028166210623e-04, 896),
]
# tests/test_solver_2d.py:178-198, problem 1 at 128^2, other cycle kinds
TESTS_2D_P1_VARIANTS = [
    # kind, max iterations (<=) or exact, rmse, rel tol, line
    ("fcycle", 8, False, 1.010315414037897e-06, 1e-9, 181),
    ("wcycle", 13, False, 1.010315424011278e-06, 1e-9, 186),
    ("fixed13", 13, True, 1.010315424011278e-06, 1e-9, 193),
    ("fixed13_fcycle", 13, False, 1.010315424011278e-06, 1e-9, 198),
]
# tests/test_solver_3d.py: (problem, mx, my, mz) -> (iterations, rmse)
TESTS_3D = [
    (1, 64, 64, 64, 19, 6.936905980033114e-05, 205),
    (1, 32, 32, 32, 19, 2.712919636528891e-04, 222),   # shadowed by a later def, never run
    (1, 32, 64, 32, 33, 2.049924560414506e-04, 239),
    (2, 64, 64, 64, 19, 6.935753745686714e-05, 288),
    (2, 32, 32, 32, 19, 2.7124687452780976e-04, 305),
    (2, 32, 64, 32, 33, 2.0495839279424837e-04, 322),
    (3, 64, 64, 64, 36, 6.93591831010889e-05, 370),
    (3, 32, 32, 32, 35, 2.712533139540335e-04, 387),
    (3, 32, 64, 32, 27, 2.4402769848607482e-04, 404),
]

ROW = re.compile(r"\(\s*(\d+)\s*,\s*(\d+)\s*\)\s*:\s*\(\s*(\d+)\s*,\s*([-+0-9.eE]+|-?np\.nan)\s*\)")


def parse_tables(path):
    text = open(path).read()
    out = {}
    for name in TABLES:
        m = re.search(name + r"\s*=\s*\{(.*?)\n\}", text, re.S)
        if m is None:
            raise RuntimeError(f"table {name} not found")
        start_line = text[:m.start()].count("\n") + 1
        rows = []
        for rm in ROW.finditer(m.group(1)):
            p, res, it, val = rm.groups()
            v = math.nan if "nan" in val else float(val)
            rows.append([int(p), int(res), int(it), v])
        out[name] = {"line": start_line, "rows": rows}
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    tables = parse_tables(os.path.join(ref, "examples", "benchmark.py"))
    data = {
        "source": "davef96/jsolver: tests/test_solver_2d.py, tests/test_solver_3d.py, examples/benchmark.py",
        "tolerance": {"rel": 1e-9, "abs": 2e-15},
        "tests_2d": [dict(problem=p, mx=mx, my=my, iterations=it, rmse=r, line=ln)
                     for (p, mx, my, it, r, ln) in TESTS_2D],
        "tests_2d_p1_variants": [dict(kind=k, iterations=it, exact=ex, rmse=r, rel=rel, line=ln)
                                 for (k, it, ex, r, rel, ln) in TESTS_2D_P1_VARIANTS],
        "tests_3d": [dict(problem=p, mx=mx, my=my, mz=mz, iterations=it, rmse=r, line=ln)
                     for (p, mx, my, mz, it, r, ln) in TESTS_3D],
        "tables_2d": tables,
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1, allow_nan=True)
    print("wrote", out, {k: len(v["rows"]) for k, v in tables.items()})


if __name__ == "__main__":
    main()
