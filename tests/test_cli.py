"""The upcgen command line and its parameters.in parser (CPU: -printpars
reads and prints the configuration without touching a GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "upcgen_amd", "upcgen")


def run(args, cwd):
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout + r.stderr


def params(text):
    out = {}
    for line in text.splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[-2].isupper():
            out[parts[-2]] = parts[-1]
    return out


def test_help_lists_reference_options(tmp_path):
    rc, out = run(["-h"], tmp_path)
    assert rc == 0
    for opt in ("-debug", "-nthreads", "-parfile"):
        assert opt in out


def test_parameters_file_parsing(tmp_path):
    (tmp_path / "p.in").write_text(
        "# comment line\n"
        "NUCLEUS_Z 54   # xenon\n"
        "NUCLEUS_A 129\n"
        "WS_R 5.49#no space before the comment\n"
        "SQRTS 5440\n"
        "PROC_ID 13\n"
        "BINS_M 77\n"
        "\n"
        "UNKNOWN_KEY 12\n"
        "FLUX_POINT 0\n"
        "BREAKUP_MODE 3\n"
        "USE_POLARIZED_CS 1\n")
    rc, out = run(["-parfile", "p.in", "-printpars"], tmp_path)
    assert rc == 0, out
    p = params(out)
    assert p["NUCLEUS_Z"] == "54" and p["NUCLEUS_A"] == "129"
    assert float(p["WS_R"]) == 5.49
    assert float(p["SQRTS"]) == 5440
    assert p["PROC_ID"] == "13" and p["BINS_M"] == "77"
    assert p["FLUX_POINT"] == "0" and p["BREAKUP_MODE"] == "3" and p["USE_POLARIZED_CS"] == "1"


def test_missing_parameters_file_uses_defaults(tmp_path):
    rc, out = run(["-parfile", "nope.in", "-printpars"], tmp_path)
    assert rc == 0
    assert "Input file not found" in out
    p = params(out)
    assert p["NUCLEUS_Z"] == "82" and p["BINS_M"] == "1001" and p["BINS_Y"] == "121"
