"""Extracts the golden vectors used by the tests from the reference's published
result files (data only; run in a container where the reference is mounted).

  gi_test_topology.json  <- benchmarks/results/main-bvh-study/leaf-size-time.csv
                            (GI_TEST rows: leaf size, topology SHA-256, nodes, depth)
"""
import csv
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent


def gi_test_topology():
    rows = []
    with open(REF / "benchmarks/results/main-bvh-study/leaf-size-time.csv") as f:
        for r in csv.DictReader(f):
            if r["sceneId"] == "GI_TEST":
                rows.append({"leafSize": int(r["leafSize"]), "topologySha256": r["topologySha256"],
                             "nodeCount": int(r["nodeCount"]), "maxDepth": int(r["maxDepth"])})
    (OUT / "gi_test_topology.json").write_text(json.dumps(rows, indent=1) + "\n")


if __name__ == "__main__":
    gi_test_topology()
