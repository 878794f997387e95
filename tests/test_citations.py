"""Every citation of the reference (`file.py:a-b`, or a bare `:a-b` after a
named file) in this repository's sources and documents points inside the
cited file: checked against the line counts of the reference checkout
(tests/golden/reference_lines.json, written by make_reference_lines.py)."""
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LINES = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_lines.json")))
BY_BASE = {}
for rel in LINES:
    BY_BASE.setdefault(os.path.basename(rel), []).append(rel)

FILE = re.compile(r"((?:[\w./-]+/)?[\w-]+\.(?:py|txt|yaml|yml|md|ipynb|xml|cfg))(?=[:\s,)`'\"]|$)")
SPAN = re.compile(r":(\d+)(?:-(\d+))?((?:,\s*:?\d+(?:-\d+)?)*)")
BARE = re.compile(r"(?<![\w./-])(?:\(|\s|,|`|^)(:\d+(?:-\d+)?)")
# At the top level, the notes checked are those that carry this line; other
# top-level documents and records (surveys, reviews, run records) are not
# this repository's sources.
MARK = "<!-- reference citations checked by tests/test_citations.py -->"
SKIP_DIRS = {"_work", "gpurun_out", "_kcache", "__pycache__", ".git", "golden", "lib", "profiles"}
TEXT = (".py", ".h", ".hip", ".cpp", ".md", ".json", "Makefile")


def _sources():
    for dp, dn, fn in os.walk(ROOT):
        dn[:] = [d for d in dn if d not in SKIP_DIRS and not d.startswith(".")]
        for f in fn:
            if dp == ROOT and not f.endswith(".py"):
                with open(os.path.join(dp, f), errors="replace") as fh:
                    if MARK not in fh.read():
                        continue
            if f.endswith(TEXT):
                yield os.path.join(dp, f)
    yield os.path.join(ROOT, "tests", "golden", "reference_kats.json")


def _candidates(name):
    if "/" in name:
        return [r for r in LINES if r == name or r.endswith("/" + name)]
    return BY_BASE.get(name, [])


def _spans(text):
    return [(int(a), int(b) if b else int(a)) for a, b in re.findall(r"(\d+)(?:-(\d+))?", text)]


def citations():
    """(path, line number, cited file candidates, [(first, last)]) of every citation."""
    out = []
    for path in _sources():
        with open(path, errors="replace") as fh:
            lines = fh.read().splitlines()
        last = None
        for ln, line in enumerate(lines, 1):
            events = []
            for m in FILE.finditer(line):
                events.append((m.start(), "file", m))
            for m in BARE.finditer(line):
                events.append((m.start(1), "bare", m))
            for pos, kind, m in sorted(events, key=lambda e: e[0]):
                if kind == "file":
                    cand = _candidates(m.group(1))
                    last = cand or None
                    s = SPAN.match(line, m.end())
                    if cand and s:
                        out.append((path, ln, cand, _spans(s.group(0))))
                elif last and not line[:m.start(1)].rstrip().endswith(tuple(".py .txt .md .xml .yaml".split())):
                    out.append((path, ln, last, _spans(m.group(1))))
    return out


def test_citations_lie_inside_the_cited_files():
    cites = citations()
    assert len(cites) > 200
    bad = []
    for path, ln, cand, spans in cites:
        if not any(all(1 <= a <= b <= LINES[c] for a, b in spans) for c in cand):
            bad.append("%s:%d cites %s lines %s (files have %s lines)" % (
                os.path.relpath(path, ROOT), ln, cand[:2], spans, [LINES[c] for c in cand[:2]]))
    assert not bad, "\n".join(bad)
