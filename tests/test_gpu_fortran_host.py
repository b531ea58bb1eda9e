"""The Fortran host (poromec_amd/fortran: iso_c_binding module pmhip and the
pm_mefdrive time-march driver) runs the Biot column on the GPU through the C
ABI; its final state and node history must match the oracle's march."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "poromec_amd" / "fortran" / "pm_mefdrive"


def test_fortran_driver_matches_oracle(tmp_path):
    if not EXE.exists():
        pytest.skip("Fortran host not built (amdflang absent)")
    from poromec_amd.fdeck import read_state, write_fdeck
    from poromec_amd.solver import SolverConfig

    deck = P.column_deck(2, 2, 6, dt=0.01)
    cfg = SolverConfig(tol=1e-10, stop_crit=1)
    nsteps = 4
    write_fdeck(tmp_path / "deck.bin", deck, cfg, nsteps, pnode=1)
    r = subprocess.run([str(EXE), str(tmp_path / "deck.bin"), str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = O.OracleModel(deck)
    hist = []
    for _ in range(nsteps):
        m.step()
        hist.append(m.u[0].copy())
    u, dp = read_state(tmp_path / "out_state.bin", m.st["nnode"])
    nv = deck.nnodev
    assert np.abs(u[:, :3] - m.u[:, :3]).max() <= 1e-9 * np.abs(m.u[:, :3]).max()
    assert np.abs(u[:nv, 3] - m.u[:nv, 3]).max() <= 1e-9 * np.abs(m.u[:nv, 3]).max()
    rows = [line.split() for line in (tmp_path / "out_up_node.txt").read_text().splitlines()
            if not line.startswith("#")]
    assert len(rows) == nsteps
    for k, row in enumerate(rows):
        vals = np.array([float(v.replace("D", "E")) for v in row[2:6]])
        assert np.allclose(vals, hist[k], rtol=1e-8, atol=1e-9 * np.abs(hist[k]).max())
