"""Writes reference_lines.json: the line count of every text file of the
mcvine/acc checkout (paths relative to it), the fixture that
tests/test_citations.py checks file:line citations against.  Run where the
checkout is at /root/reference: python tests/golden/make_reference_lines.py"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SUFFIXES = (".py", ".txt", ".yaml", ".yml", ".md", ".xml", ".cfg", ".ipynb")
out = {}
for dp, dn, fn in os.walk(REF):
    dn[:] = [d for d in dn if not d.startswith(".")]
    for f in fn:
        if f.endswith(SUFFIXES):
            p = os.path.join(dp, f)
            with open(p, errors="replace") as fh:
                out[os.path.relpath(p, REF)] = sum(1 for _ in fh)
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_lines.json"), "w") as fh:
    json.dump(dict(sorted(out.items())), fh, indent=0)
print(len(out), "files")
