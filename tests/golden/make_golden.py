"""Writes tests/golden/kats.json: known-answer vectors taken from the reference's
own tests (inputs and the expected outputs those tests assert), plus the
published SplitMix64 output stream. Run: python tests/golden/make_golden.py

Sources (reference repository paths):
* exact order with index tie-break: crates/polariseq-core/src/doublets.rs:1702-1709
* exact path, 4 cells in 2-D: crates/polariseq-core/src/neighbors.rs:248-265
* self excluded: neighbors.rs:300-313; counts: neighbors.rs:285-297; cosine: neighbors.rs:316-332
* sq_dist/dot tail lengths and vectors: crates/polariseq-core/src/simd.rs:94-119
* tiny inputs: crates/polariseq-core/src/nndescent.rs:1183-1199 (n=5, k=10 -> 4 per row)
* SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference C code), seed 0, first four
  outputs. nndescent.rs:80-87's splitmix(x) is the output for state x, so output i
  is splitmix(i * 0x9E3779B97F4A7C15).
"""

import json
from pathlib import Path

KATS = {
    "line_points": {
        "source": "crates/polariseq-core/src/doublets.rs:1702-1709",
        "points": [[0.0], [1.0], [2.0], [3.0], [10.0]],
        "k": 2,
        "expect_rows": {"0": [1, 2], "1": [0, 2], "4": [3, 2]},
    },
    "four_cells": {
        "source": "crates/polariseq-core/src/neighbors.rs:248-265",
        "points": [[0.0, 0.0], [1.0, 0.0], [10.0, 10.0], [0.0, 1.0]],
        "n_neighbors": 3,
        "expect_row0_set": [1, 3],
        "expect_row0_dist": 1.0,
        "tol": 1e-5,
    },
    "self_excluded": {
        "source": "crates/polariseq-core/src/neighbors.rs:300-313",
        "points": [[0.0, 0.0], [1.0, 0.0], [2.0, 0.0]],
        "n_neighbors": 2,
    },
    "counts": {
        "source": "crates/polariseq-core/src/neighbors.rs:285-297",
        "n": 10, "d": 3, "n_neighbors": 5, "expect_per_row": 4,
        "points_rule": "x[i] = i * 0.1 for the flattened 10x3 array",
    },
    "cosine": {
        "source": "crates/polariseq-core/src/neighbors.rs:316-332",
        "points": [[1.0, 0.0], [2.0, 0.0], [-1.0, 0.0]],
        "n_neighbors": 2,
        "expect_row0_first": 1,
        "expect_row0_dist_below": 1e-5,
    },
    "simd_tails": {
        "source": "crates/polariseq-core/src/simd.rs:94-119",
        "lengths": [0, 1, 7, 8, 9, 15, 16, 17, 50, 64, 101],
        "a_rule": "((i*7) % 13) * 0.25 - 1.0",
        "b_rule": "((i*5) % 11) * 0.5 - 2.0",
        "rel_tol": 1e-3,
    },
    "tiny_inputs": {
        "source": "crates/polariseq-core/src/nndescent.rs:1183-1199",
        "n": 5, "d": 3, "k": 10, "expect_per_row": 4,
    },
    "splitmix64_seed0": {
        "source": "SplitMix64 published output stream, seed 0",
        "outputs": ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4", "0x06c45d188009454f",
                    "0xf88bb8a8724c81ec"],
    },
}

if __name__ == "__main__":
    out = Path(__file__).with_name("kats.json")
    out.write_text(json.dumps(KATS, indent=2) + "\n")
    print("wrote", out)
