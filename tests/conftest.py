import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session", autouse=True)
def _built():
    """Build the CPU oracle if it is missing (cheap); the HIP library is
    built by __graft_entry__.build() beforehand."""
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True)
    ref_src = "/root/reference/kokkos"
    ref_lib = os.path.join(ROOT, "oracle", "_ref", "libref_mesh.so")
    if os.path.isdir(ref_src) and not os.path.exists(ref_lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    yield


def read_inp(path):
    """The nine values of a miniaero.inp (Options.h:91-99)."""
    t = open(path).read().split()
    return dict(problem_type=int(t[0]), lx=float(t[1]), ly=float(t[2]), lz=float(t[3]),
                angle=float(t[4]), nx=int(t[5]), ny=int(t[6]), nz=int(t[7]), ntimesteps=int(t[8]),
                dt=float(t[9]), output_results=int(t[10]), output_frequency=int(t[11]),
                second_order_space=int(t[12]), viscous=int(t[13]))


# The reference's integration cases: (directory, ranks, rel_tol, floor)
# tolerances from the runner scripts (tests/*/ *_test.sh of the reference).
GOLD_CASES = [
    ("3D_Sod_Serial", 1, 1e-3, 1e-6),
    ("3D_Sod_Parallel", 4, 1e-3, 1e-6),
    ("Ramp_Serial", 1, 1e-2, 1e-3),
    ("Ramp_Parallel", 16, 1e-2, 1e-3),
    ("FlatPlate_Serial", 1, 1e-2, 1e-3),
    ("FlatPlate_Parallel", 8, 1e-2, 1e-3),
]


def gold_rows(case, ranks, coords):
    """The gold result rows of a case in single-domain cell order (parallel
    golds re-assembled by cell-centroid coordinates)."""
    import numpy as np

    from miniaero_amd.results import assemble_by_coordinates

    d = os.path.join(GOLDEN, case)
    if ranks == 1:
        return np.loadtxt(os.path.join(d, "results.gold"), ndmin=2)
    parts = [np.loadtxt(os.path.join(d, f"results.{r}.gold"), ndmin=2) for r in range(ranks)]
    return assemble_by_coordinates(parts, coords)


def rows_text(rows):
    return "\n".join("".join("%g\t" % v for v in r) for r in rows) + "\n"
