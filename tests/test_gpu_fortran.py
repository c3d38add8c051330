"""The Fortran host (quickpic-opensource_amd/fortran): its class layer over
the C ABI drives the GPU slice solver through the reference's method sequence
(simulation_class.f03:611-736), and through the fused qp_slice_run.  Both are
compared with the oracle's restatement of the same slices (FP64, 1e-10
relative).

The driver runs as a child process, before this process touches the GPU.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import qpo

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "quickpic-opensource_amd", "fortran", "bin", "qp_slice_driver")
IND, INDZ, NSL = 5, 5, 4


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


@pytest.fixture(scope="module")
def oracle_state():
    grid = qpo.Grid(IND, IND, INDZ)
    bspec = qpo.BeamSpec(center=(0.0, 0.0, 1.0), sigma=(1.0, 1.0, 0.6), np_=(16, 16, 32),
                         peak=4.0)
    osim = qpo.Simulation(grid, [qpo.SpeciesSpec(ppc=(2, 2))], [bspec], dt=10.0, iter_=2)
    osim.begin_step()
    for j in range(1, NSL + 1):
        osim.slice(j)
    return osim


def run_driver(mode):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C quickpic-opensource_amd/fortran)")
    out = os.path.join(tempfile.mkdtemp(), f"fields_{mode}.bin")
    print(f"[gpu-test] fortran driver mode {mode}", flush=True)
    r = subprocess.run([DRIVER, str(IND), str(INDZ), str(NSL), "0", str(mode), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = (1 << IND) + 1
    raw = np.fromfile(out, dtype=np.float64)
    sizes = [1 * n * n, 2 * n * n, 3 * n * n, 3 * n * n, 2 * n * n]
    assert raw.size == sum(sizes)
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    # Fortran f(dim, ld1, ld2) column-major -> (dim, ld2, ld1)
    real = [p.reshape(n, n, d).transpose(2, 0, 1) for p, d in zip(parts[:3], (1, 2, 3))]
    kf = [p.reshape(n, n, d).transpose(2, 0, 1) for p, d in zip(parts[3:], (3, 2))]
    return real + kf


@pytest.mark.parametrize("mode", [0, 1])
def test_fortran_driver_matches_oracle(oracle_state, mode):
    psit, fxy, bxyz, cu_k, epw_k = run_driver(mode)
    o = oracle_state
    errs = {"psit": rel(psit, o.psit), "fxy": rel(fxy, o.fxy), "bxyz": rel(bxyz, o.bxyz),
            "cu_k": rel(cu_k, o.cu_k), "epw_k": rel(epw_k, o.epw_k)}
    print(f"mode {mode}: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    assert all(v < 1e-10 for v in errs.values()), errs
