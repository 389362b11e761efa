"""Write reference_kat.json: the inputs and expected outputs held by the
reference's own known-answer tests for the hot path (data only).

Sources (schnelLFMM):
  src/bed.rs:511-562   round trip n=7 (odd padding), p=3, eigenstrat closed form
  src/bed.rs:564-582   missing code byte 0b11_10_01_00 -> [-sqrt2, 0, 0, sqrt2]
  src/bed.rs:584-611   norm modes byte 0b00_11_10_00, n=3
  src/precompute.rs:97-157  4x2 X, lambda=1
  src/testing.rs:1071-1075  median vectors
  src/testing.rs:1078-1113  safe_inv SPD and singular (eps = 1e-8 * 4)
  src/testing.rs:1116-1126  histogram t=[1..5] -> GIF ~ 9/0.4549 (+-0.1)
  src/testing.rs:1129-1159  calibrate ordering t=[3,-2], gif 1
  tests/integration.rs:754-778  subset_indices vectors
  src/nmf.rs:530-585   dense NMF 4x3 (with/without missing), K=2, error < 1
Run: python tests/golden/make_reference_kat.py
"""
import json, math, os

S2 = math.sqrt(2.0)
kat = {
    "decode_roundtrip": {
        "n": 7, "p": 3, "norm": "eigenstrat", "tol": 1e-10,
        "genotypes": [[0, 1, 2], [2, 0, 1], [1, 1, 0], [0, 2, 2], [2, 1, 0], [1, 0, 1], [0, 0, 2]],
    },
    "missing_code": {"packed": [0b11100100], "n": 4, "norm": "eigenstrat",
                     "expected": [-S2, 0.0, 0.0, S2], "tol": 1e-10},
    "norm_modes": {"packed": [0b00111000], "n": 3,
                   "center_only": [-1.0, 0.0, 1.0],
                   "eigenstrat": [-S2, 0.0, S2], "tol": 1e-10},
    "precompute": {"x": [[1.0, 0.5], [0.3, -0.2], [-0.5, 1.0], [0.8, 0.1]], "lambda": 1.0, "tol": 1e-10},
    "median": [{"v": [1.0, 2.0, 3.0], "m": 2.0}, {"v": [1.0, 2.0, 3.0, 4.0], "m": 2.5}, {"v": [5.0], "m": 5.0}],
    "safe_inv": {"spd": [[2.0, 1.0], [1.0, 3.0]], "spd_tol": 1e-10,
                 "singular": [[1.0, 2.0], [2.0, 4.0]], "singular_eps": 4e-8, "singular_tol": 1e-6},
    "histogram": {"t": [1.0, 2.0, 3.0, 4.0, 5.0], "gif": 9.0 / 0.4549, "tol": 0.1},
    "calibrate_order": {"t": [3.0, -2.0], "gif_sqrt": 1.0},
    "subset_indices": [
        {"spec": "all", "n": 10, "idx": list(range(10))},
        {"spec": 0.5, "n": 10, "idx": [0, 2, 4, 6, 8]},
        {"spec": 0.25, "n": 10, "idx": [0, 4, 8]},
        {"spec": 1.0, "n": 10, "idx": list(range(10))},
        {"spec": 0.1, "n": 25, "idx": [0, 10, 20]},
        {"spec": [3, 7, 11], "n": 100, "idx": [3, 7, 11]},
    ],
    "nmf_dense": {"y": [[3.0, 1.0, 2.0], [2.0, 0.0, 1.0], [0.0, 1.0, 0.0], [1.0, 2.0, 1.0]],
                  "y_missing": [[3.0, 1.0, 2.0], [2.0, None, 1.0], [0.0, 1.0, 0.0], [1.0, 2.0, None]],
                  "k": 2, "n_iter": 200, "n_iter_missing": 300, "eps": 1e-16, "max_err": 1.0},
}
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kat.json")
with open(out, "w") as f:
    json.dump(kat, f, indent=1)
print("wrote", out)
