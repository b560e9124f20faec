"""Extract the golden vectors the oracle is pinned against from the reference's data files.

Reads (text only, nothing executed):
  benchmarks/tensorkit_semantic_oracle.out   (TensorKit 0.17 symbol tables,
      invariant stream, planar repartition fixtures; sha256 pinned in
      benchmarks/tensorkit_semantic_oracle.sha256)
  benchmarks/qspace_su2_oracle.out
and writes tests/golden/tensorkit_oracle.json.  Known-answer vectors that live
inside the reference's Rust tests are transcribed by hand into
tests/golden/kat.json with their file:line.

Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""
import hashlib
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse(text):
    out = {"F_SU2": [], "R_SU2": [], "dim_SU2": [], "stream": {}, "repartition": {}}
    section = None
    stream = None
    rep = None
    for line in text.splitlines():
        if line.startswith("== section"):
            section = line
            continue
        if line.startswith("-- ") and line.endswith(" --"):
            name = line[3:-3]
            if "section 3" in (section or ""):
                stream = name
                out["stream"].setdefault(stream, {})
            elif "section 4" in (section or ""):
                rep = name
            continue
        m = re.match(r"F SU2 (\d+) (\d+) (\d+) (\d+) (\d+) (\d+) = (\S+)", line)
        if m:
            out["F_SU2"].append([int(x) for x in m.groups()[:6]] + [float(m.group(7))])
            continue
        m = re.match(r"R SU2 (\d+) (\d+) (\d+) = (\S+)", line)
        if m:
            out["R_SU2"].append([int(x) for x in m.groups()[:3]] + [float(m.group(4))])
            continue
        m = re.match(r"dim SU2 (\d+) = (\S+)  twist = (\S+)  fs = (\S+)", line)
        if m:
            out["dim_SU2"].append([int(m.group(1)), float(m.group(2)), float(m.group(3)),
                                   float(m.group(4))])
            continue
        if "section 3" in (section or "") and stream in ("U1", "SU2", "fZ2xU1xSU2"):
            m = re.match(r"(s\d+[ab]?) norm = (\S+)  tr = (\S+)", line)
            if m:
                out["stream"][stream][m.group(1)] = [float(m.group(2)), float(m.group(3))]
                continue
            m = re.match(r"s8 tr\(H\*H\) = (\S+)", line)
            if m:
                out["stream"][stream]["s8"] = float(m.group(1))
                continue
        if "section 4" in (section or "") and rep:
            m = re.match(r"data = \[(.*)\]", line)
            if m:
                out["repartition"][rep] = [float(x) for x in m.group(1).split(",")]
                continue
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    path = os.path.join(ref, "benchmarks", "tensorkit_semantic_oracle.out")
    with open(path, "rb") as fh:
        raw = fh.read()
    data = parse(raw.decode())
    data["source"] = "benchmarks/tensorkit_semantic_oracle.out"
    data["sha256"] = hashlib.sha256(raw).hexdigest()
    with open(os.path.join(ref, "benchmarks", "qspace_su2_oracle.out")) as fh:
        data["qspace_su2"] = fh.read().splitlines()
    with open(os.path.join(HERE, "tensorkit_oracle.json"), "w") as fh:
        json.dump(data, fh, indent=1)
    print(f"F={len(data['F_SU2'])} R={len(data['R_SU2'])} streams={list(data['stream'])} "
          f"repartition={list(data['repartition'])}")


if __name__ == "__main__":
    main()
