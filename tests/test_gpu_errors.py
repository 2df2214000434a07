"""Error returns of the pressure iteration and a regression case of source's level-0 terms, on the GPU.

psolve's failures are returned as sw_status codes where the reference calls exit(1):
* divergence, errp > 1 after the first iteration (src/nhswps/psolve.rs:349-352): a seeded k^-3 state
  with rms zeta = 2f and k0 = 4, on which the oracle (the reference algorithm) diverges at its second
  iteration with errp ~1.6e3 (128x4) and ~1.2e2 (96x6), far from the threshold;
* a non-finite residual (SW_ERR_NONFINITE; the reference does not test for it: NaN > toler is false,
  so its loop stops and returns the NaN fields): a NaN in qs.
Both on the device-controlled iteration (128x4: wave-per-sequence transforms, nz with a compile-time
specialisation) and on the host-controlled one (96x6).  No kernel faults or hangs here: a NaN flows
through the transforms as a value and the stopping rule reads it.

source's level-0 Jacobian J(ff, z) when it reuses psolve's derivatives (advance's correctors, and its
predictor when ggen is set): ff = deal(ri dpn) vanishes at level 0 when dpn does there, as it does in
every reference state (psolve writes dpn's levels 1..nz only), so a state with a non-zero dpn at
level 0 checks that the z derivatives at level 0 are the ones formed for it."""
import numpy as np
import pytest
from conftest import gold, rel_l2
from _tolerances import STEP_TOL

pytestmark = pytest.mark.gpu
ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]


def seeded(sw, oracle, ng, nz, seed, k0, zeta_frac):
    qs, ds = sw.synth_kminus3(ng, nz, seed=seed, k0=k0, zeta_rms=zeta_frac * 4 * np.pi)
    o = oracle.State(oracle.Spectral(ng, nz), qs=qs, ds=ds, gs=np.zeros_like(qs), ggen=True)
    o.spectral.main_invert(o.qs, o.ds, o.gs, o.r, o.u, o.v, o.zeta)
    g = sw.State(sw.Spectral(ng, nz), ggen=True, **{f: getattr(o, f) for f in ALL})
    return g, o


@pytest.mark.parametrize("ng,nz", [(128, 4), (96, 6)])
def test_psolve_diverged(sw, oracle, ng, nz):
    g, o = seeded(sw, oracle, ng, nz, seed=5, k0=4.0, zeta_frac=2.0)
    rc, it, errp = oracle.psolve(o)
    assert rc == -3 and it == 2 and errp > 100.0, (rc, it, errp)
    with pytest.raises(sw.SwError) as e:
        sw.psolve(g)
    assert e.value.code == -3, str(e.value)


@pytest.mark.parametrize("ng,nz", [(128, 4), (96, 6)])
def test_psolve_nonfinite(sw, oracle, ng, nz):
    g, _ = seeded(sw, oracle, ng, nz, seed=5, k0=8.0, zeta_frac=0.5)
    qs = g.get("qs")
    qs[ng * ng + 3 * ng + 5] = np.nan
    s = g.spectral
    r, u, v, zeta = s.field(), s.field(), s.field(), s.field()
    s.main_invert(qs, g.get("ds"), g.get("gs"), r, u, v, zeta)
    assert np.isnan(r).any()
    for name, a in (("qs", qs), ("r", r), ("u", u), ("v", v), ("zeta", zeta)):
        g.set(name, a)
    with pytest.raises(sw.SwError) as e:
        sw.psolve(g)
    assert e.value.code == -7, str(e.value)


def with_dpn_level0(fields, ng, seed):
    rng = np.random.default_rng(seed)
    f = dict(fields)
    dpn = f["dpn"].copy()
    dpn[:ng * ng] = 1e-3 * rng.standard_normal(ng * ng)
    f["dpn"] = dpn
    return f


def test_advance_nonzero_dpn_level0_golden_state(sw, oracle):
    ng, nz, t0 = 24, 4, 0.624_999_999_999_999_9
    p = f"nhswps/testdata/advance/{ng}_{nz}/"
    f = with_dpn_level0({n: gold(p + n + ".bin") for n in ALL}, ng, 7)
    g = sw.State(sw.Spectral(ng, nz), t=t0, ggen=True, **f)
    o = oracle.State(oracle.Spectral(ng, nz), t=t0, ggen=True, **f)
    _, iters = sw.advance(g)
    rc, _, iters2 = oracle.advance(o)
    assert rc == 0 and iters == iters2
    errs = {n: rel_l2(g.get(n), getattr(o, n)) for n in ALL}
    print("24x4 with dpn level 0:", {n: f"{e:.1e}" for n, e in errs.items()})
    for n in ALL:
        assert errs[n] < STEP_TOL[n], n


def test_advance_nonzero_dpn_level0_wave_path(sw, oracle):
    ng, nz = 128, 4
    g0, o0 = seeded(sw, oracle, ng, nz, seed=9, k0=8.0, zeta_frac=0.5)
    f = with_dpn_level0({n: getattr(o0, n) for n in ALL}, ng, 8)
    g = sw.State(sw.Spectral(ng, nz), ggen=True, **f)
    o = oracle.State(oracle.Spectral(ng, nz), ggen=True, **f)
    _, iters = sw.advance(g)
    rc, _, iters2 = oracle.advance(o)
    assert rc == 0 and iters == iters2
    errs = {n: rel_l2(g.get(n), getattr(o, n)) for n in ALL}
    print("128x4 with dpn level 0:", {n: f"{e:.1e}" for n, e in errs.items()})
    for n in ALL:
        assert errs[n] < STEP_TOL[n], n
