"""Extracts the golden vectors used by the tests from the reference's published
result files (data only; run in a container where the reference is mounted).

  gi_test_topology.json  <- benchmarks/results/main-bvh-study/leaf-size-time.csv
                            (GI_TEST rows: leaf size, topology SHA-256, nodes, depth)
  catalog_topology.json  <- the same file's CITY_OF_NIGHT_V1 and SIGNAL_CHAMBER rows
  hero_topology.json     <- benchmarks/results/main-bvh-study/hero-sensitivity.csv
                            (Layered Hero: scale, leaf size, lambda, topology SHA-256)
                            and core8-detail.csv (uniform / weighted rows at leaf 8)
  hero_counters.json     <- benchmarks/results/main-bvh-study/core8-detail.csv
                            (per-frame counters of the uniform and weighted rows of
                            the Hero scales: 1920x1080, 1 spp, depth 12, stack 32;
                            and the weighted-vs-uniform kernel-time reduction)
  overlap_10k.json       <- the same file's overlap-10k uniform and weighted rows
  bvh_matrix.json        <- bvh-node-count-matrix.csv, bvh-depth-matrix.csv,
                            bvh-bytes-matrix.csv (uniform and weighted columns of
                            the Hero and overlap rows: nodes, depth, packed bytes)
  modeled_work.json      <- core8-detail.csv, every single-layout row (all eight
                            methods): the six test counters and medianModeledWork
"""
import csv
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent
STUDY = REF / "benchmarks/results/main-bvh-study"
SCALES = {"hero-96": 96, "hero-1k": 1_000, "hero-10k": 10_000, "hero-100k": 100_000, "hero-1m": 1_000_000}
COUNTERS = {"rays": "medianRays", "aabb_tests": "medianAabbTests", "primitive_tests": "medianPrimitiveTests",
            "sphere_tests": "medianSphereTests", "box_tests": "medianBoxTests", "plane_tests": "medianPlaneTests",
            "affine_sphere_tests": "medianAffineSphereTests", "affine_box_tests": "medianAffineBoxTests",
            "internal_node_visits": "medianInternalNodeVisits", "leaf_node_visits": "medianLeafNodeVisits"}


def gi_test_topology():
    rows = []
    with open(STUDY / "leaf-size-time.csv") as f:
        for r in csv.DictReader(f):
            if r["sceneId"] == "GI_TEST":
                rows.append({"leafSize": int(r["leafSize"]), "topologySha256": r["topologySha256"],
                             "nodeCount": int(r["nodeCount"]), "maxDepth": int(r["maxDepth"])})
    (OUT / "gi_test_topology.json").write_text(json.dumps(rows, indent=1) + "\n")


def catalog_topology():
    rows = []
    with open(STUDY / "leaf-size-time.csv") as f:
        for r in csv.DictReader(f):
            if r["sceneId"] in ("CITY_OF_NIGHT_V1", "SIGNAL_CHAMBER"):
                rows.append({"sceneId": r["sceneId"], "leafSize": int(r["leafSize"]),
                             "topologySha256": r["topologySha256"], "nodeCount": int(r["nodeCount"]),
                             "maxDepth": int(r["maxDepth"])})
    (OUT / "catalog_topology.json").write_text(json.dumps(rows, indent=1) + "\n")


def _core8_rows(ids):
    with open(STUDY / "core8-detail.csv") as f:
        return [r for r in csv.DictReader(f) if r["rowId"] in ids and r["methodId"] in ("uniform", "weighted")]


def _core8_hero_rows():
    return _core8_rows(SCALES)


def hero_topology():
    cells = []
    with open(STUDY / "hero-sensitivity.csv") as f:
        for r in csv.DictReader(f):
            cells.append({"rowId": r["rowId"], "scale": int(r["scale"]), "leafSize": int(r["leafSize"]),
                          "lambda": float(r["lambda"]), "topologySha256": r["topologySha256"]})
    for r in _core8_hero_rows():
        cells.append({"rowId": r["rowId"], "scale": SCALES[r["rowId"]], "leafSize": 8,
                      "lambda": 0.0 if r["methodId"] == "uniform" else 1.0, "topologySha256": r["topologySha256"]})
    (OUT / "hero_topology.json").write_text(json.dumps(cells, indent=0) + "\n")


def _counter_row(r, scale):
    return {"rowId": r["rowId"], "rowOrdinal": int(r["rowOrdinal"]), "scale": scale,
            "method": r["methodId"], "layoutCount": int(r["layoutCount"]),
            "topologySha256": r["topologySha256"],
            "timeReductionPercent": float(r["timeReductionPercent"]),
            "counters": {k: int(r[c]) for k, c in COUNTERS.items()}}


def hero_counters():
    rows = [_counter_row(r, SCALES[r["rowId"]]) for r in _core8_hero_rows()]
    (OUT / "hero_counters.json").write_text(json.dumps(rows, indent=1) + "\n")


def overlap_rows():
    rows = [_counter_row(r, 10_000) for r in _core8_rows({"overlap-10k"})]
    (OUT / "overlap_10k.json").write_text(json.dumps(rows, indent=1) + "\n")


def bvh_matrix():
    tables = {}
    for key, name in (("nodeCount", "bvh-node-count-matrix.csv"), ("maxDepth", "bvh-depth-matrix.csv"),
                      ("bytes", "bvh-bytes-matrix.csv")):
        with open(STUDY / name) as f:
            tables[key] = {r["rowId"]: r for r in csv.DictReader(f)}
    rows = []
    random_rows = [f"F{f}-{n}" for f in range(1, 5) for n in ("100", "1k", "10k")]
    for row_id in list(SCALES) + ["overlap-10k"] + random_rows:
        for method in ("uniform", "weighted"):
            # random rows hold medians over 100 layouts (a half-integer when the middle pair differs)
            rows.append({"rowId": row_id, "method": method,
                         **{k: float(t[row_id][method]) if row_id in random_rows else int(t[row_id][method])
                            for k, t in tables.items()}})
    (OUT / "bvh_matrix.json").write_text(json.dumps(rows, indent=0) + "\n")


def modeled_work():
    tests = ("medianAabbTests", "medianSphereTests", "medianBoxTests", "medianPlaneTests",
             "medianAffineSphereTests", "medianAffineBoxTests")
    rows = []
    with open(STUDY / "core8-detail.csv") as f:
        for r in csv.DictReader(f):
            if r["layoutCount"] == "1":
                rows.append({"rowId": r["rowId"], "method": r["methodId"],
                             **{k: int(r[k]) for k in tests}, "medianModeledWork": r["medianModeledWork"]})
    (OUT / "modeled_work.json").write_text(json.dumps(rows, indent=0) + "\n")


def random_reductions():
    """Weighted-vs-uniform medians over the 100 layouts of each random row (core8-detail.csv)."""
    rows = []
    with open(STUDY / "core8-detail.csv") as f:
        for r in csv.DictReader(f):
            if r["rowId"].startswith("F") and r["methodId"] == "weighted":
                rows.append({"rowId": r["rowId"], "layoutCount": int(r["layoutCount"]),
                             "timeReductionPercent": float(r["timeReductionPercent"]),
                             "workReductionPercent": float(r["workReductionPercent"]),
                             "medianRays": float(r["medianRays"]), "medianModeledWork": float(r["medianModeledWork"])})
    (OUT / "random_reductions.json").write_text(json.dumps(rows, indent=0) + "\n")


if __name__ == "__main__":
    gi_test_topology()
    catalog_topology()
    hero_topology()
    hero_counters()
    overlap_rows()
    bvh_matrix()
    modeled_work()
    random_reductions()
