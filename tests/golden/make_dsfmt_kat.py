"""Extract the dSFMT-19937 known-answer vectors that ship with the reference
(+MCmatlab/src/dSFMT-src-2.2.3/dSFMT.19937.out.txt, the generator's own
published check file) into tests/golden/dsfmt19937_kat.json.

Data only: for each `init_gen_rand(seed)` section, the seed, the interval and
the printed values (15 decimals).  Run in a container that has the reference:
    python tests/golden/make_dsfmt_kat.py
"""
import json
import os
import re
import sys

SRC = "/root/reference/+MCmatlab/src/dSFMT-src-2.2.3/dSFMT.19937.out.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dsfmt19937_kat.json")


def main(src=SRC, out=OUT):
    sections = []
    cur = None
    with open(src) as f:
        for line in f:
            m = re.match(r"init_gen_rand\((\d+)\) generated randoms (\S+, \S+)", line)
            if m:
                cur = {"seed": int(m.group(1)), "interval": m.group(2), "values": []}
                sections.append(cur)
                continue
            if line.startswith("init_by_array"):
                cur = None
                continue
            if cur is not None and re.match(r"^[0-9.\s]+$", line) and line.strip():
                cur["values"].extend(float(v) for v in line.split())
    with open(out, "w") as f:
        json.dump({"source": "dSFMT.19937.out.txt (dSFMT 2.2.3 known-answer file)",
                   "idstring": "dSFMT2-19937:117-19:ffafffffffb3f-ffdfffc90fffd",
                   "sections": sections}, f, indent=0)
    print(f"wrote {out}: {len(sections)} sections, {sum(len(s['values']) for s in sections)} values")


if __name__ == "__main__":
    main(*sys.argv[1:])
