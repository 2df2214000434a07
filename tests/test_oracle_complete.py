"""The oracle's time loop, diagnostics and savegrid output against the reference's end-to-end case
(src/test.rs:27-36, 86-260): ng = 32, nz = 4, the default parameters (time step 1/32, save interval
0.25), run to t = 1.0 from the balanced strip that the reference's balinit writes
(testdata/balinit/32_4_sw_init.r8: q, delta, gamma as three records of an 8-byte header and ng^2
little-endian f64 values).  swto3d copies each record to the nz + 1 levels (src/swto3d.rs:10-39),
and the nhswps loader transforms them, zeroes the mean and de-aliases (src/nhswps/mod.rs:159-277).

Pinned here: monitor.asc (diagnose, one row per step, src/nhswps/mod.rs:666-711) and ecomp.asc
(savegrid's energies, :365-420) token for token, as the reference's own test compares them, and the
.r4 field dumps (:491-652: 3-D ql, d, g, r, w, pn; vertically integrated d, zeta, q, g, h) at the
reference's 1e-8 (src/utils.rs:62-79).  Closed forms pin diagnose on its own: a uniform flow gives
umax = |U|, and a uniform vorticity c on every level gives zrms = zmax = |c|."""
import os

import numpy as np
import pytest
from conftest import GOLD

CASE = os.path.join(GOLD, "testdata")
NG, NZ = 32, 4
N3 = ["ql", "d", "g", "r", "w", "pn"]
N2 = ["d", "zeta", "q", "g", "h"]


def sw_init(path, ng):
    """The three 2-D records of an sw_init.r8 file (an 8-byte header before each)."""
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == 3 * (ng * ng + 1)
    return raw.reshape(3, ng * ng + 1)[:, 1:]


def loaded_state(oracle, ng, nz, recs):
    """swto3d's layer copy, then the nhswps loader: ptospc3d, zero (0, 0) mode, times filt."""
    s = oracle.Spectral(ng, nz)
    filt = s.get("filt")
    n2 = ng * ng
    fields = {}
    for name, rec in zip(("qs", "ds", "gs"), recs):
        phys = np.tile(rec, nz + 1)
        spec = s.field()
        s.ptospc3d(phys, spec, 0, nz)
        for iz in range(nz + 1):
            spec[iz * n2] = 0.0
            spec[iz * n2:(iz + 1) * n2] *= filt
        fields[name] = spec
    return oracle.State(s, **fields)


def monitor_tokens(rows):
    """diagnose's line format: {:.5} {:.6} {:.6} {:.6} {:.6}."""
    return [tok for r in rows for tok in ["%.5f" % r[0]] + ["%.6f" % x for x in r[1:]]]


def ecomp_tokens(rows):
    """savegrid's ecomp line: {:.6} {:.9} {:.9} {:.9} {:.9} {:.9}."""
    return [tok for r in rows for tok in ["%.6f" % r[0]] + ["%.9f" % x for x in r[1:]]]


def read_r4(path, n):
    """Records of an f32 time then n f32 values (append_output, src/nhswps/mod.rs:654-663)."""
    a = np.fromfile(path, dtype="<f4")
    assert a.size % (n + 1) == 0
    a = a.reshape(-1, n + 1)
    return a[:, 0], a[:, 1:]


@pytest.fixture(scope="module")
def recs():
    return sw_init(os.path.join(CASE, "balinit", "32_4_sw_init.r8"), NG)


def test_complete_32_4_monitor_and_ecomp(oracle, recs):
    st = loaded_state(oracle, NG, NZ, recs)
    rc, mon, eco, _ = oracle.run(st, 1.0, 0.25)
    assert rc == 0 and len(mon) == 33 and len(eco) == 5
    ref_m = open(os.path.join(CASE, "complete", "32_4", "monitor.asc")).read().split()
    ref_e = open(os.path.join(CASE, "complete", "32_4", "ecomp.asc")).read().split()
    assert monitor_tokens(mon) == ref_m
    assert ecomp_tokens(eco) == ref_e


def test_complete_32_4_field_dumps(oracle, recs):
    b = loaded_state(oracle, NG, NZ, recs)
    dt, tsim, ngsave = 1.0 / NG, 1.0, 8
    sp = b.spectral
    sp.main_invert(b.qs, b.ds, b.gs, b.r, b.u, b.v, b.zeta)
    saves = []

    def save():
        sp.main_invert(b.qs, b.ds, b.gs, b.r, b.u, b.v, b.zeta)
        assert oracle.psolve(b)[0] == 0
        saves.append((np.float32(b.t),) + oracle.savegrid(b))

    while b.t <= tsim:
        if int(round(b.t / dt)) % ngsave == 0:
            save()
            b.ggen = False
        else:
            b.ggen = True
        assert oracle.advance(b)[0] == 0
    if int(b.t / dt) % ngsave == 0:
        save()
    assert len(saves) == 5
    n2 = NG * NG
    for i, name in enumerate(N3):
        t, ref = read_r4(os.path.join(CASE, "complete", "32_4", "3d", name + ".r4"), n2 * (NZ + 1))
        mine = np.array([s[1][i] for s in saves])
        assert np.array_equal(t, np.array([s[0] for s in saves])), name
        np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-8, err_msg="3d/" + name)
    for i, name in enumerate(N2):
        _, ref = read_r4(os.path.join(CASE, "complete", "32_4", "2d", name + ".r4"), n2)
        mine = np.array([s[2][i] for s in saves])
        np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-8, err_msg="2d/" + name)


def test_diagnose_closed_forms(oracle):
    ng, nz = 24, 4
    n = ng * ng * (nz + 1)
    s = oracle.Spectral(ng, nz)
    st = oracle.State(s, u=np.full(n, 0.3), v=np.full(n, -0.4), zeta=np.full(n, -1.25), t=0.5)
    d = oracle.diagnose(st)
    assert d[0] == 0.5
    assert abs(d[4] - 0.5) < 1e-15  # umax = |(0.3, -0.4)|
    assert abs(d[2] - 1.25) < 1e-14 and d[3] == 1.25  # zrms = zmax = |c|
    assert abs(d[1] - 0.5 * 1.25 ** 2) < 1e-14
    # zrms weights the end levels by 1/2 (trapezoid): vorticity on the top level alone gives zrms^2 = c^2/(2 nz)
    z = np.zeros(n)
    z[nz * ng * ng:] = 2.0
    d = oracle.diagnose(oracle.State(s, zeta=z))
    assert abs(d[2] ** 2 - 4.0 / (2 * nz)) < 1e-14 and d[3] == 2.0 and d[4] == 0.0
