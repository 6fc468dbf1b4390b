"""Write MBM fixtures from the reference's problem files (data only).

    python tests/golden/make_mbm_subset.py /path/to/reference

From scripts/panda_problems.json and scripts/fetch_problems.json (same JSON
layout, original 1-based indices kept in "index"):

* mbm_<robot>_subset.json: the first two valid problems of every family.
  The planner tests read it: the reference marks these starts and goals
  valid (collision free), a known answer for the pose checker;
* mbm_<robot>_valid.json.gz: every valid problem (699 Panda, 679 Fetch),
  compact JSON, gzip-compressed: the full-dataset run of the benchmark
  (the reference harness's scripts/evaluate_mbm.cpp:119-152 loop).
"""
import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    for robot in ("panda", "fetch"):
        with open(os.path.join(ref, "scripts", f"{robot}_problems.json")) as f:
            data = json.load(f)
        out = {"robot": data["robot"], "joints": data["joints"], "problems": {}}
        for family, pset in data["problems"].items():
            out["problems"][family] = [p for p in pset if p.get("valid")][:2]
        with open(os.path.join(HERE, f"mbm_{robot}_subset.json"), "w") as f:
            json.dump(out, f, indent=0)
        print(robot, {k: [p["index"] for p in v] for k, v in out["problems"].items()})
        full = {"robot": data["robot"], "joints": data["joints"],
                "problems": {family: [p for p in pset if p.get("valid")] for family, pset in data["problems"].items()}}
        raw = json.dumps(full, separators=(",", ":")).encode()
        # mtime 0: the same input gives the same bytes
        with open(os.path.join(HERE, f"mbm_{robot}_valid.json.gz"), "wb") as f:
            f.write(gzip.compress(raw, 9, mtime=0))
        print(robot, sum(len(v) for v in full["problems"].values()), "valid problems")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
