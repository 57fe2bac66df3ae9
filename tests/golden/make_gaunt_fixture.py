"""Extract the golden Gaunt values of HelFEM's src/general/gaunt_test.cpp.

The reference test hard-codes 552 (call, expected value) pairs for
gaunt_coefficient(L,M,l,m,lp,mp) and modified_gaunt_coefficient(lj,mj,L,M,li,mi).
This script reads that file as text and writes only the numbers (inputs and
expected outputs) to gaunt_reference.json.  It runs in the build container,
where the reference tree is mounted; the GPU host only sees the JSON.
"""
import json
import re
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/general/gaunt_test.cpp")
OUT = Path(__file__).with_name("gaunt_reference.json")

call = re.compile(r"val=helfem::gaunt::(gaunt_coefficient|modified_gaunt_coefficient)\(([-0-9, ]+)\);")
ref = re.compile(r"ref=([-+0-9.eE]+);")

rows = []
pending = None
for line in SRC.read_text().splitlines():
    m = call.search(line)
    if m:
        pending = (m.group(1), [int(v) for v in m.group(2).split(",")])
        continue
    r = ref.search(line)
    if r and pending is not None:
        kind = "coeff" if pending[0] == "gaunt_coefficient" else "mod_coeff"
        rows.append({"kind": kind, "args": pending[1], "value": float(r.group(1))})
        pending = None

OUT.write_text(json.dumps({"source": "src/general/gaunt_test.cpp", "entries": rows}))
print(f"wrote {len(rows)} entries to {OUT}")
