"""Copies the reference's own test inputs and expected outputs for this path
into tests/golden/ as fixtures (data only).  Run in a container where the
reference checkout is mounted; the GPU tests only read the copies."""
import json
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

INPUTS = ["tests/inputs/sailors.csv", "tests/inputs/reserves.csv", "tests/inputs/boats.csv"]
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
}


def main():
    for f in INPUTS:
        shutil.copy(os.path.join(REF, f), os.path.join(HERE, os.path.basename(f)))
    gold = {}
    for name, f in OUTPUTS.items():
        with open(os.path.join(REF, f)) as fh:
            gold[name] = {"source": f, "expected": fh.read().strip()}
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as fh:
        json.dump(gold, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
