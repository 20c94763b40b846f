"""results.<rank> writer/reader and a Python 3 counterpart of the reference's
numeric_text_diff comparator.

Format (TimeSolverExplicitRK4.h:520-537): one row per owned cell in cell
order, `x y z rho rhou rhov rhow rhoE`, each value followed by a tab, default
ostream precision (6 significant digits, %g), newline.

numeric_text_diff (tests/tools/numeric_text_diff:1-106 of the reference, a
Python 2 script) compares two such files line by line: identical text passes,
otherwise every field pair must satisfy |a-b| / max(|a|,|b|) <= rel_tol unless
both |a| and |b| are <= floor.  Defaults rel_tol = 1e-3, floor = 1e-6.
"""
from __future__ import annotations

import numpy as np


def format_rows(coords: np.ndarray, U: np.ndarray) -> str:
    rows = []
    for x, u in zip(np.asarray(coords), np.asarray(U)):
        rows.append("".join("%g\t" % v for v in (*x, *u)))
    return "\n".join(rows) + ("\n" if rows else "")


def write_results(path: str, coords: np.ndarray, U: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write(format_rows(coords, U))


def read_results(path: str) -> np.ndarray:
    return np.loadtxt(path, ndmin=2)


def numeric_text_diff(text1: str, text2: str, rel_tol: float = 1e-3, floor: float = 1e-6):
    """Returns (same, message).  Lines are compared in order; a differing
    line count fails."""
    l1 = text1.splitlines()
    l2 = text2.splitlines()
    if len(l1) != len(l2):
        return False, "Files have differing line counts."
    diffs = []
    for a_line, b_line in zip(l1, l2):
        if a_line == b_line:
            continue
        fa, fb = a_line.split(), b_line.split()
        if len(fa) != len(fb):
            return False, "Number of fields differ for the following line:\n   %s   %s" % (a_line, b_line)
        for x, y in zip(fa, fb):
            if x == y:
                continue
            try:
                xv, yv = float(x), float(y)
            except ValueError:
                diffs.append("-%s\n+%s" % (a_line, b_line))
                break
            if abs(xv) > floor or abs(yv) > floor:
                if abs(xv - yv) / max(abs(xv), abs(yv)) > rel_tol:
                    diffs.append("-%s\n+%s" % (a_line, b_line))
                    break
    if diffs:
        return False, "Files differ with relative_tol=%g & floor=%g:\n\n%s" % (rel_tol, floor, "\n".join(diffs))
    return True, "Files are the same."


def numeric_file_diff(path1: str, path2: str, rel_tol: float = 1e-3, floor: float = 1e-6):
    with open(path1) as f1, open(path2) as f2:
        return numeric_text_diff(f1.read(), f2.read(), rel_tol, floor)


def assemble_by_coordinates(parts: list[np.ndarray], coords: np.ndarray, decimals: int = 5) -> np.ndarray:
    """Re-assembles per-rank result rows into the single-domain cell order by
    matching the printed cell-centroid coordinates."""
    allrows = np.vstack(parts)
    key = {tuple(np.round(r[:3], decimals)): r for r in allrows}
    out = np.empty((coords.shape[0], allrows.shape[1]))
    for i, x in enumerate(coords):
        out[i] = key[tuple(np.round([float("%g" % v) for v in x], decimals))]
    return out
