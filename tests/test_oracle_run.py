"""CPU pins of the oracle's energy diagnostic and time loop (src/nhswps/mod.rs:285-345, 365-420).

energy: closed forms -- a uniform flow u = U over the whole column (r = v = w = 0) has
ekin = (1/2) (2 pi)^2 U^2 (the trapezoid weights sum to nz levels of thickness HBAR/nz), and a top
surface z = HBAR (1 + eps cos x) has epot = pi^2 c^2 eps^2.
run: the save cadence and ggen flags against the same loop driven step by step through the oracle's
advance, main_invert, psolve and energy."""
import numpy as np
import pytest
from conftest import gold

HBAR, CSQ = 0.4, (2 * np.pi) ** 2
ALL = ["u", "v", "w", "z", "zx", "zy", "r", "ri", "aa", "zeta", "pn", "dpn", "ps", "qs", "ds", "gs"]


def test_energy_closed_forms(oracle):
    ng, nz, U, eps = 24, 4, 0.7, 0.01
    s = oracle.Spectral(ng, nz)
    n2 = ng * ng
    x = 2 * np.pi * np.arange(ng) / ng - np.pi
    z = np.zeros(n2 * (nz + 1))
    z[nz * n2:] = HBAR * (1 + eps * np.repeat(np.cos(x), ng))  # element (i=y, j=x) at i + ng j
    st = oracle.State(s, u=np.full(n2 * (nz + 1), U), z=z)
    e = oracle.energy(st)
    assert e[1] == 0.0 and e[2] == e[3]
    assert abs(e[2] - 0.5 * (2 * np.pi) ** 2 * U ** 2) < 1e-12 * e[2]
    assert abs(e[4] - np.pi ** 2 * CSQ * eps ** 2) < 1e-12 * e[4]
    assert abs(e[5] - e[2] - e[4]) < 1e-15 * e[5]


def test_run_matches_stepwise_loop(oracle):
    ng, nz = 24, 4
    p = "nhswps/testdata/advance/24_4/"
    f = {n: gold(p + n + ".bin") for n in ALL}
    dt = 1.0 / ng
    a = oracle.State(oracle.Spectral(ng, nz), **f)
    b = oracle.State(oracle.Spectral(ng, nz), **f)
    tsim, tgsave = 6.5 * dt, 3 * dt
    rc, mon, eco, its = oracle.run(a, tsim, tgsave)
    # saves at itime 0, 3, 6, and after the loop: t is then a rounding below 7 dt, and the final test
    # truncates t / dt (mod.rs:335-336) to 6
    assert rc == 0 and len(mon) == 7 and len(eco) == 4
    # the same loop, step by step
    sp = b.spectral
    sp.main_invert(b.qs, b.ds, b.gs, b.r, b.u, b.v, b.zeta)
    rows, erows = [], []
    while b.t <= tsim:
        itime = int(round(b.t / dt))
        if itime % 3 == 0:
            sp.main_invert(b.qs, b.ds, b.gs, b.r, b.u, b.v, b.zeta)
            assert oracle.psolve(b)[0] == 0
            erows.append(oracle.energy(b))
            b.ggen = False
        else:
            b.ggen = True
        rc2, d, _ = oracle.advance(b)
        assert rc2 == 0
        rows.append(d)
    if int(b.t / dt) % 3 == 0:
        sp.main_invert(b.qs, b.ds, b.gs, b.r, b.u, b.v, b.zeta)
        assert oracle.psolve(b)[0] == 0
        erows.append(oracle.energy(b))
    np.testing.assert_array_equal(mon, np.array(rows))
    np.testing.assert_array_equal(eco, np.array(erows))
    for n in ALL:
        np.testing.assert_array_equal(getattr(a, n), getattr(b, n))
