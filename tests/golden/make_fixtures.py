"""Extract the reference's own known-answer vectors into JSON fixtures.

Run in a container that has the reference checked out (default
/root/reference).  It copies data only -- the Philox4x32 known-answer lines and
the u01fixedpt<float> IEEE vectors that Random123 1.09 ships with its tests:

  Random123/examples/kat_vectors              (philox4x32, R = 7 and 10)
  Random123/examples/old_kat_vectors          (philox4x32, R = 7 and 10)
  Random123/examples/ut_uniform_IEEEkatvectors.hpp  (u01fixedpt<float> on u32)

Random123 is Copyright 2010-2013 D. E. Shaw Research, BSD-3-Clause; the
vectors keep that attribution in the fixture headers.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
EX = os.path.join(REF, "Random123", "examples")
OUT = os.path.dirname(os.path.abspath(__file__))
ATTR = "Random123 1.09 test vectors, (c) D. E. Shaw Research, BSD-3-Clause"


def philox_lines(path):
    rows = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) == 12 and tok[0] == "philox4x32":
                rounds = int(tok[1])
                ctr = [int(t, 16) for t in tok[2:6]]
                key = [int(t, 16) for t in tok[6:8]]
                exp = [int(t, 16) for t in tok[8:12]]
                rows.append({"rounds": rounds, "ctr": ctr, "key": key, "expected": exp})
    return rows


def u01_vectors(path):
    # insert("u01fixedpt<float>(<base> + <type>(a)) a=<a>", <value>L);
    pat = re.compile(r'insert\("u01fixedpt<float>\((minu32|maxu32|mini32|maxi32) ([+-]) '
                     r'u?int32_t\(a\)\) a=(\d+)", ([-0-9.e+]+)L\);')
    base = {"minu32": 0, "maxu32": 0xFFFFFFFF, "mini32": 0x80000000, "maxi32": 0x7FFFFFFF}
    rows = []
    with open(path) as f:
        for line in f:
            m = pat.search(line)
            if not m:
                continue
            b, sign, a, val = m.group(1), m.group(2), int(m.group(3)), m.group(4)
            r = (base[b] + a) if sign == "+" else (base[b] - a)
            rows.append({"r": r & 0xFFFFFFFF, "expected": val})
    return rows


def main():
    ph = philox_lines(os.path.join(EX, "kat_vectors")) + \
        philox_lines(os.path.join(EX, "old_kat_vectors"))
    u01 = u01_vectors(os.path.join(EX, "ut_uniform_IEEEkatvectors.hpp"))
    with open(os.path.join(OUT, "philox4x32_kat.json"), "w") as f:
        json.dump({"attribution": ATTR, "vectors": ph}, f, indent=0)
    with open(os.path.join(OUT, "u01fixedpt_f32_kat.json"), "w") as f:
        json.dump({"attribution": ATTR, "vectors": u01}, f, indent=0)
    print(len(ph), "philox4x32 vectors;", len(u01), "u01fixedpt<float> vectors")


if __name__ == "__main__":
    main()
