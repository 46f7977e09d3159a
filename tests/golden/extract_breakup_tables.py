"""Writes tests/golden/breakup_tables.npz: the numeric photo-absorption data
of upcgen's calcBreakupProb, read as text from the reference source file.

calcBreakupProb (src/UpcCrossSection.cpp:752-1019 of the upcgen tree) builds
its energy / cross-section tables from literal arrays (e1, s1, e2, s2, e3, s3,
sen, sigt, sigtn, sa; lines 778-850) and a few literal constants of the
lead branch (si1, g1, o1, o0; the Regge parameters x, y, eps, eta, em; the
grid step delo).  This script parses those literals out of the reference
file so the tests can compare the oracle's tables with numbers that were not
typed by hand.  It reads the file as text only; nothing of the reference is
compiled, imported or run.  The fixture holds numbers (data), not source.

Run in the build container (the reference is not on the GPU machines):
    python tests/golden/extract_breakup_tables.py [path/to/UpcCrossSection.cpp]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/src/UpcCrossSection.cpp"
ARRAYS = ("e1", "s1", "e2", "s2", "e3", "s3", "sa", "sen", "sigt", "sigtn")
NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def _function_body(text, name):
    start = text.index(f"UpcCrossSection::{name}(")
    open_brace = text.index("{", start)
    depth = 0
    for i in range(open_brace, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[open_brace:i + 1]
    raise ValueError(f"unbalanced braces in {name}")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _array(body, name):
    m = re.search(r"\b" + name + r"\s*\[(\d+)\]\s*=\s*\{([^}]*)\}", body)
    if m is None:
        raise ValueError(f"array {name} not found")
    declared = int(m.group(1))
    vals = [float(v) for v in re.findall(NUM, m.group(2))]
    return declared, np.array(vals)


def _scalar(text, name):
    m = re.search(r"\b" + name + r"\s*=\s*(" + NUM + r")\s*;", text)
    if m is None:
        raise ValueError(f"constant {name} not found")
    return float(m.group(1))


def extract(src=DEFAULT_SRC):
    text = open(src).read()
    body = _strip_comments(_function_body(text, "calcBreakupProb"))
    out = {}
    for name in ARRAYS:
        declared, vals = _array(body, name)
        out[name] = vals
        out[name + "_declared"] = np.int64(declared)
    # the lead branch of the first-call initialisation: from its "zp = 82;"
    # (the last assignment of that value; the first is the declaration) to
    # the loop that copies sen into sa
    last = list(re.finditer(r"\bzp\s*=\s*82\s*;", body))[-1]
    rest = body[last.end():]
    lead = rest[:re.search(r"for\s*\(", rest).start()]
    for name in ("si1", "g1", "o1", "o0"):
        out["pb_" + name] = np.float64(_scalar(lead, name))
    for name in ("delo", "omax1n", "x", "y", "eps", "eta", "em", "hbarcmev", "pi"):
        out[name] = np.float64(_scalar(body, name))
    out["alpha_inv"] = np.float64(float(re.search(r"/\s*(" + NUM + r")\s*;\s*", body[body.index("zcon ="):]).group(1)))
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    data = extract(src)
    np.savez(os.path.join(HERE, "breakup_tables.npz"), **data)
    print({k: (v.shape if np.ndim(v) else float(v)) for k, v in data.items()})
