"""Copies the reference's own test inputs, physical plans and expected
outputs for this path into tests/golden/ as fixtures (data only), keeping the
reference's "inputs/..." layout so the plans run unchanged with
tests/golden as their base directory.  Run in a container where the reference
checkout is mounted; the tests only read the copies."""
import json
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

INPUTS = ["tests/inputs/sailors.csv", "tests/inputs/reserves.csv", "tests/inputs/boats.csv",
          "tests/inputs/A.csv", "tests/inputs/B.csv",
          "tests/inputs/parser/A.csv", "tests/inputs/parser/B.csv", "tests/inputs/parser/catalog.json",
          "tests/inputs/plans/catalog.json",
          "tests/inputs/synthetic/100-30cols-fixed.field1",
          "tests/inputs/synthetic/100-30cols-fixed.field4",
          "tests/inputs/tpch/part_0.001.tbl", "tests/inputs/tpch/partsupp_0.001.tbl"]
# physical plans of tests/olap/tests-plan-parsing.cpp and the GPU plan fixtures
PLANS = ["reduce-scan", "reduce-twofields-scan", "reduce-join", "reduce-select",
         "reduce-join-record", "reduce-join-record-nonselective", "reduce-scan-bin",
         "reduce-scan-gpu", "gpu-group-by", "reduce-join-parts", "reduce-scan-part",
         "reduce-scan-partsupp"]
# the Star Schema Benchmark's queries as bare physical plans over prepared
# ("block") relations: tests/inputs/plans/proteus_bare_plans/q*.sql_bare.json
SSB_BARE = ["q1_1", "q1_2", "q1_3", "q2_1", "q2_2", "q2_3", "q3_1", "q3_2", "q3_3", "q3_4",
            "q4_1", "q4_2", "q4_3"]
OUTPUTS = {
    # RelBuilder operator tests, core/olap/tests/operator-tests/tests-join-sailors.cpp
    "sailorsScan": "tests/olap/outputs/tests-sailors/sailorsScan.csv",
    "sailorsSel": "tests/olap/outputs/tests-sailors/sailorsSel.csv",
    "sailorsScanBoats": "tests/olap/outputs/tests-sailors/sailorsScanBoats.csv",
    "sailorsJoinLeft3": "tests/olap/outputs/tests-sailors/sailorsJoinLeft3.csv",
    # tests/olap/tests-output.cpp
    "3wayJoin": "tests/olap/outputs/tests-output/3wayJoin.json",
    # tests/olap/tests-plan-parsing.cpp (PlanTest.Join / Scan / Select)
    "reduce-join": "tests/olap/outputs/tests-plan-parsing/reduce-join-log.json",
    "reduce-scan": "tests/olap/outputs/tests-plan-parsing/reduce-scan-log.json",
    "reduce-select": "tests/olap/outputs/tests-plan-parsing/reduce-select-log.json",
    "reduce-twofields-scan": "tests/olap/outputs/tests-plan-parsing/reduce-twofields-scan-log.json",
    "reduce-join-record": "tests/olap/outputs/tests-plan-parsing/reduce-join-record-log.json",
    "reduce-join-record-nonselective":
        "tests/olap/outputs/tests-plan-parsing/reduce-join-record-nonselective-log.json",
    "reduce-scan-bin": "tests/olap/outputs/tests-plan-parsing/reduce-scan-bin-log.json",
}


def main():
    for f in INPUTS:
        dst = os.path.join(HERE, f[len("tests/"):])
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copy(os.path.join(REF, f), dst)
    for pl in PLANS:
        dst = os.path.join(HERE, "inputs", "plans", pl + ".json")
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copy(os.path.join(REF, "tests/inputs/plans", pl + ".json"), dst)
    for q in SSB_BARE:
        dst = os.path.join(HERE, "inputs", "plans", "ssb", q + ".json")
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        with open(os.path.join(REF, "tests/inputs/plans/proteus_bare_plans", q + ".sql_bare.json")) as fh:
            plan = json.load(fh)
        with open(dst, "w") as fh:  # the same document without its indentation
            json.dump(plan, fh, separators=(",", ":"))
    gold = {}
    for name, f in OUTPUTS.items():
        with open(os.path.join(REF, f)) as fh:
            gold[name] = {"source": f, "expected": fh.read().strip()}
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as fh:
        json.dump(gold, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
