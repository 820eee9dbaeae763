"""Writes scan_kat.json: the known-answer vectors of the reference's scan
tests, kept as data.

  SphWave2D/ParallelScan.cpp:124-155   ParallelScanTest: x = {1,0,1,0,1,2,1,2,
      1,2,0,1,0,2,1,0}, truth = serial exclusive scan, "Failed" on mismatch
  UniformGrid2D/ParallelScan.cpp:113-126  sixteen ones (no comparison in the
      reference; the exclusive scan of ones is 0..15)
"""
import json
from pathlib import Path

x = [1, 0, 1, 0, 1, 2, 1, 2, 1, 2, 0, 1, 0, 2, 1, 0]
kat = {
    "sphwave2d_parallelscantest": {
        "input": x,
        "expected_exclusive": [0, 1, 1, 2, 2, 3, 5, 6, 8, 9, 11, 11, 12, 12, 14, 15],
        "source": "SphWave2D/ParallelScan.cpp:124-155",
    },
    "uniformgrid2d_ones": {
        "input": [1] * 16,
        "expected_exclusive": list(range(16)),
        "source": "UniformGrid2D/ParallelScan.cpp:113-126",
    },
}
Path(__file__).with_name("scan_kat.json").write_text(json.dumps(kat, indent=1) + "\n")
