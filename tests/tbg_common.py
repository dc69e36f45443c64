"""Shared helpers for the twisted-bilayer fixtures (test_tbg.jl of the reference)."""
import json
import os

import numpy as np

FIXTURES = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tbg_fixtures.json")))


def index(label: str, n: int) -> int:
    """0-based index of a 1-based fixture label such as "14", "end", "end-1", "end/2+1"."""
    if label.startswith("end/2"):
        base = n // 2
        return base + (int(label[5:]) if len(label) > 5 else 0) - 1
    if label.startswith("end"):
        return n - 1 + (int(label[3:]) if len(label) > 3 else 0)
    return int(label) - 1


def check(ev: np.ndarray, table: dict, tol: float) -> None:
    for label, val in table.items():
        i = index(label, len(ev))
        assert abs(ev[i] - val) < tol, (label, ev[i], val)
