"""Collect the reference's golden error-norm files into JSON fixtures.

Reads (only where /root/reference exists):
  samples/validation/data_ref/validation_3d_<Center>_<bc6>_typeGreen=<k>.txt
      lines "<nglob0> <L2> <Linf> [<L2> <Linf>]*"   (validation_3d.cpp:430-464)
  samples/solve_vtube/data_ref/vtube_<Center>_<order>_<dir><symx><symy>_typeGreen=<k>.txt
and writes validation_ref.json / vtube_ref.json beside this script. The JSON
holds the same numbers as the text files (data, not code).
Run:  python tests/golden/make_fixtures.py
"""
import json
import os
import re

REF = "/root/reference/samples"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = []
    d = os.path.join(REF, "validation", "data_ref")
    for f in sorted(os.listdir(d)):
        m = re.match(r"validation_3d_(\w+?)_(\d{6})_typeGreen=(\d+)\.txt$", f)
        if not m:
            continue
        for line in open(os.path.join(d, f)):
            p = line.split()
            if p:
                out.append({"center": m.group(1), "bc": m.group(2), "kernel": int(m.group(3)),
                            "N": int(p[0]), "err": [float(x) for x in p[1:]]})
    json.dump(out, open(os.path.join(HERE, "validation_ref.json"), "w"), indent=0)
    vt = []
    d = os.path.join(REF, "solve_vtube", "data_ref")
    for f in sorted(os.listdir(d)):
        m = re.match(r"vtube_(\w+?)_(\d+)_(-?\d)(-?\d)(-?\d)_typeGreen=(\d+)\.txt$", f)
        if not m:
            continue
        for line in open(os.path.join(d, f)):
            p = line.split()
            if p:
                vt.append({"center": m.group(1), "order": int(m.group(2)), "dir": int(m.group(3)),
                           "symx": int(m.group(4)), "symy": int(m.group(5)), "kernel": int(m.group(6)),
                           "N": int(p[0]), "err": [float(x) for x in p[1:]]})
    json.dump(vt, open(os.path.join(HERE, "vtube_ref.json"), "w"), indent=0)
    print(len(out), "validation lines,", len(vt), "vtube lines")


if __name__ == "__main__":
    main()
