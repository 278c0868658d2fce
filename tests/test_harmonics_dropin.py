"""The drop-in's host harmonics R, S, dR, dS (include/solidfmm/harmonics.hpp)
against the reference: one C++ program over the public API
(oracle/harmonics_check.cpp) built against both libraries, compared
coefficient by coefficient.  CPU only; skipped where the programs were not
built (oracle/Makefile target dropin)."""
import os
import subprocess

import numpy as np
import pytest

REF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def _run(name):
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built")
    r = subprocess.run([path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def test_harmonics_match_reference_program():
    ref_lines = _run("ref_harmonics_check")
    mine = _run("dropin_harmonics_check")
    assert len(ref_lines) == len(mine) > 0
    kinds = set()
    for a, b in zip(ref_lines, mine):
        ka, kb = a.split()[:4], b.split()[:4]
        assert ka == kb
        va = np.array([float(t) for t in a.split()[4:]])
        vb = np.array([float(t) for t in b.split()[4:]])
        # double: rounding; float: the drop-in evaluates in double and rounds
        # once, the reference accumulates float rounding along the recurrences
        tol = 1e-14 if ka[3] == "8" else 1e-5
        assert np.abs(va - vb).max() <= tol * max(np.abs(va).max(), 1e-300), a[:40]
        kinds.add((ka[0], ka[3]))
    assert kinds == {(k, t) for k in ("R", "S", "dR", "dS") for t in ("4", "8")}
