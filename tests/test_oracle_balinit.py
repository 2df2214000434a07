"""The oracle's balinit (src/balinit.rs:21-436) against the reference's own end-to-end cases
(src/balinit.rs:441-481): qq_init.r8 in, sw_init.r8 out, at 18x2 and 32x4, with the reference's
tolerance (abs 1e-13 per value, src/utils.rs:41-59).  The two cases pin every step of the
restatement: the height and vorticity inversion, the velocity, the mean flow, each iteration's
balanced gamma and delta, and the stopping rule (the iteration counts differ: 8 and 9)."""
import os

import numpy as np
import pytest
from conftest import GOLD

CASE = os.path.join(GOLD, "testdata", "balinit")


def read_case(ng, nz):
    q = np.fromfile(os.path.join(CASE, f"{ng}_{nz}_qq_init.r8"), dtype="<f8")
    assert q.size == ng * ng + 1
    out = np.fromfile(os.path.join(CASE, f"{ng}_{nz}_sw_init.r8"), dtype="<f8").reshape(3, ng * ng + 1)
    assert not out[:, 0].any()  # each record's header is a zero
    return q[1:].copy(), out[:, 1:]


@pytest.mark.parametrize("ng,nz,iters", [(18, 2, 8), (32, 4, 9)])
def test_oracle_balinit_matches_reference(oracle, ng, nz, iters):
    q0, ref = read_case(ng, nz)
    qq, dd, gg, it = oracle.Spectral(ng, nz).balinit(q0)
    assert it == iters
    for name, a, b in zip(("qq", "dd", "gg"), (qq, dd, gg), ref):
        assert np.abs(a - b).max() <= 1e-13, (name, np.abs(a - b).max())


def test_oracle_balinit_zero_anomaly(oracle):
    """A zero PV anomaly is already balanced: every output is exactly zero after one iteration (whose
    change, 0 / (0 + small), is below the tolerance)."""
    qq, dd, gg, it = oracle.Spectral(32, 4).balinit(np.zeros(32 * 32))
    assert it == 1
    assert not qq.any() and not dd.any() and not gg.any()
