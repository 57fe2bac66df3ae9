"""Extract Legendre P_l^m / Q_l^m reference values for x > 1.

HelFEM's src/legendre/legendre_reference_data.h holds Maple values (35
digits) of P_l^m(x) and Q_l^m(x) in HelFEM's convention.  The prolate-
spheroidal two-electron kernel evaluates them at x = cosh(mu) >= 1, so this
keeps the x > 1 entries with m <= 8, rounded to doubles, as legendre_reference.json.
"""
import json
import re
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/legendre/legendre_reference_data.h")
OUT = Path(__file__).with_name("legendre_reference.json")
row = re.compile(r'\{\s*(\d+),\s*(\d+),\s*"([^"]+)",\s*"([^"]+)",\s*"([^"]+)"\}')
rows = []
for m in row.finditer(SRC.read_text()):
    l, mm, x, P, Q = int(m.group(1)), int(m.group(2)), float(m.group(3)), float(m.group(4)), float(m.group(5))
    if x > 1.0 and mm <= 8:
        rows.append([l, mm, x, P, Q])
OUT.write_text(json.dumps({"source": "src/legendre/legendre_reference_data.h", "columns": ["l", "m", "x", "P", "Q"],
                           "rows": rows}))
print(f"wrote {len(rows)} rows; l range {min(r[0] for r in rows)}..{max(r[0] for r in rows)}; "
      f"x values {sorted(set(r[2] for r in rows))}")
