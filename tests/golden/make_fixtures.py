"""Writes tests/golden/reference_fixtures.json from the data held in the reference's source text.

Run once in a container where the reference is mounted (it is not needed afterwards):
    python tests/golden/make_fixtures.py /root/reference
Extracted (as data, by regular expression over the text):
  * fpm     -- the 30 target parameters, frg.cu:149-178
  * gamma   -- frg.cu:41
  * u, u2   -- the two 23-entry patch vectors left in comments, frg.cu:73-95 and :99-121
  * tolerances -- the #defines of htpytrack.cu:9-31, frgkinematics.cu:7-8, frgsupport.cu:7-8
"""
import json
import os
import re
import sys

NUM = r"(-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?)"


def main(ref: str):
    frg = open(os.path.join(ref, "frg.cu")).read()
    fpm = [None] * 30
    for m in re.finditer(r"fpm\[0\]\.vals\[(\d+)\]\s*=\s*\{\s*" + NUM + r"\s*,\s*" + NUM + r"\s*\}", frg):
        fpm[int(m.group(1))] = [float(m.group(2)), float(m.group(3))]
    g = re.search(r"Cplx gamma\s*=\s*\{\s*" + NUM + r"\s*,\s*" + NUM + r"\s*\}", frg)
    patches = {}
    for name in ("u", "u2"):
        vals = [None] * 23
        for m in re.finditer(r"//\s*" + name + r"\.vals\[(\d+)\]\s*=\s*\{\s*" + NUM + r"\s*,\s*" + NUM + r"\s*\}", frg):
            vals[int(m.group(1))] = [float(m.group(2)), float(m.group(3))]
        patches[name] = vals
    defs = {}
    for fn in ("htpytrack.cu", "frgkinematics.cu", "frgsupport.cu"):
        txt = open(os.path.join(ref, fn)).read()
        for m in re.finditer(r"#define\s+(\w+)\s+" + NUM, txt):
            defs[f"{fn}:{m.group(1)}"] = float(m.group(2))
    out = {"fpm": fpm, "gamma": [float(g.group(1)), float(g.group(2))], "u": patches["u"], "u2": patches["u2"],
           "defines": defs}
    assert all(v is not None for v in fpm) and all(v is not None for v in patches["u"])
    assert all(v is not None for v in patches["u2"])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
