"""Collect the reference energies that pin the Fock-build path.

Reads HelFEM's tests/refs/ci.json (recorded energies, commit 65b2ed4) and
tests/cases.json (the command-line arguments of each case) and writes the
numbers for the cases this engine reproduces to reference_energies.json.
Runs in the build container, where the reference tree is mounted.
"""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).with_name("reference_energies.json")
WANT = ["atomic-H-hf-u", "atomic-He-hf-r", "atomic-Be-hf-r", "atomic-He-lda-r", "atomic-Be-lda-r",
        "atomic-Ne-hf-r", "atomic-Ne-lda-r", "diatomic-H2-hf-r", "diatomic-He-dummy-hf", "diatomic-H2-purem-on",
        "diatomic-Ne-dummy-hf", "diatomic-Ne-dummy-lda", "atomic-He-lda-u", "atomic-Be-lda-u",
        "diatomic-NH-hf-u", "diatomic-NH-hf-absm", "diatomic-NH-lda-u", "diatomic-NH-lda-absm", "diatomic-H2-purem-off",
        "atomic-He-gga-r", "atomic-He-gga-u", "diatomic-Ne-dummy-gga", "atomic-He-mgga-r", "atomic-He-mgga-u",
        "diatomic-Ne-dummy-mgga", "atomic-He-hybrid-r", "atomic-He-hybrid-u",
        "atomic-H-rsh-erfc-u", "atomic-H-rsh-yukawa-u", "atomic-He-hf-u", "atomic-Be-hf-u", "diatomic-H2-hf-u",
        "atomic-O-3P-p0-hf", "atomic-O-3P-pplus-hf", "atomic-O-3P-pminus-hf", "diatomic-CH-piplus-hf",
        "diatomic-Ne-dummy-hf-absm", "diatomic-H2-absm-hf", "atomic-Be-nop-l0", "atomic-Be-nop-pinned",
        "diatomic-CH-absm-hf", "diatomic-CH-lda-absm"]

refs = json.loads((REF / "tests/refs/ci.json").read_text())["cases"]
cases = {c["name"]: c for c in json.loads((REF / "tests/cases.json").read_text())["cases"]}
out = {}
for name in WANT:
    entry = {"args": cases[name]["args"] if name in cases else None}
    if name in refs:
        entry.update({k: v for k, v in refs[name].items() if not k.startswith("_") and k not in ("err",)})
    out[name] = entry
# tests/README.md:79-80 records the atomic Ne HF limit directly
out["atomic-Ne-hf-r"]["total_readme"] = -128.5470981094
OUT.write_text(json.dumps(out, indent=1, sort_keys=True))
print(f"wrote {len(out)} cases to {OUT}")
