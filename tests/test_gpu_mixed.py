"""Accuracy of the mixed-precision step (SW_MIXED: psolve's coefficient fields and its iteration in FP32,
with the pressure tridiagonal and the residual in FP64; the state and every other stage in FP64)
against FP64, per field after one step: q (qs), delta (ds), gamma (gs), h (z at the top level) and p
(pn).  Against the oracle at 128x8 and against the FP64 GPU step at 512x16, from the benchmark's seeded
k^-3 state, and at 1536x24 (three-wave transforms) from the large workload's state.  Every field must be within the 1e-5 bar, and within a regression bound of about twice the
error measured on MI355X; each test prints its errors."""
import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu
ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]
OUT = {"q": "qs", "delta": "ds", "gamma": "gs", "p": "pn"}


def seeded_fields(sw, oracle, ng, nz, seed=1):
    qs, ds = sw.synth_kminus3(ng, nz, seed=seed, k0=8.0, zeta_rms=0.5 * 4 * np.pi)
    o = oracle.State(oracle.Spectral(ng, nz), qs=qs, ds=ds, gs=np.zeros_like(qs), ggen=True)
    o.spectral.main_invert(o.qs, o.ds, o.gs, o.r, o.u, o.v, o.zeta)
    return o, {f: getattr(o, f).copy() for f in ALL}


def field_errors(get_a, get_b, ng, nz):
    err = {k: rel_l2(get_a(f), get_b(f)) for k, f in OUT.items()}
    n2 = ng * ng
    err["h"] = rel_l2(get_a("z")[nz * n2:], get_b("z")[nz * n2:])
    return err


def test_mixed_one_step_vs_oracle_128x8(sw, oracle):
    ng, nz = 128, 8
    o, f = seeded_fields(sw, oracle, ng, nz)
    m = sw.State(sw.Spectral(ng, nz, sw.MIXED), ggen=True, **f)
    sw.advance(m)
    rc, _, _ = oracle.advance(o)
    assert rc == 0
    err = field_errors(m.get, lambda n: getattr(o, n), ng, nz)
    print(f"mixed {ng}x{nz} one step vs oracle:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    # measured: q 1.3e-9, delta 2.9e-7, gamma 4.9e-7, h 8.2e-11, p 2.3e-7
    bound = {"q": 3e-9, "delta": 6e-7, "gamma": 1e-6, "h": 2e-10, "p": 5e-7}
    for k, v in err.items():
        assert v < min(bound[k], 1e-5), (k, v)


@pytest.mark.timeout(300)
def test_mixed_one_step_vs_fp64_512x16(sw, oracle):
    ng, nz = 512, 16
    _, f = seeded_fields(sw, oracle, ng, nz)
    d = sw.State(sw.Spectral(ng, nz, sw.FP64), ggen=True, **f)
    m = sw.State(sw.Spectral(ng, nz, sw.MIXED), ggen=True, **f)
    sw.advance(d)
    sw.advance(m)
    err = field_errors(m.get, d.get, ng, nz)
    print(f"mixed {ng}x{nz} one step vs FP64 GPU:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    # measured: q 1.2e-9, delta 2.9e-7, gamma 7.4e-7, h 5.2e-12, p 1.0e-7
    bound = {"q": 3e-9, "delta": 6e-7, "gamma": 1.5e-6, "h": 1.2e-11, "p": 2.5e-7}
    for k, v in err.items():
        assert v < min(bound[k], 1e-5), (k, v)


@pytest.mark.timeout(300)
def test_mixed_one_step_vs_fp64_1536x24(sw):
    """The large mixed-radix workload (ng = 1536 on three-wave transforms, nz = 24): seeded k^-3 PV (rms
    zeta = f/2) plus an independent random delta of rms 0.01 f, inverted by the FP64 GPU main_invert."""
    ng, nz = 1536, 24
    qs, ds = sw.synth_kminus3(ng, nz, seed=1, k0=8.0, zeta_rms=0.5 * 4 * np.pi, delta_rms=0.01 * 4 * np.pi)
    gs = np.zeros_like(qs)
    s64 = sw.Spectral(ng, nz, sw.FP64)
    r, u, v, zeta = s64.field(), s64.field(), s64.field(), s64.field()
    s64.main_invert(qs, ds, gs, r, u, v, zeta)
    f = dict(qs=qs, ds=ds, gs=gs, r=r, u=u, v=v, zeta=zeta)
    d = sw.State(s64, ggen=True, **f)
    sw.advance(d)
    df = {n: d.get(n) for n in ("qs", "ds", "gs", "z", "pn")}
    del d
    m = sw.State(sw.Spectral(ng, nz, sw.MIXED), ggen=True, **f)
    sw.advance(m)
    err = field_errors(m.get, lambda n: df[n], ng, nz)
    print(f"mixed {ng}x{nz} one step vs FP64 GPU:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    # measured: q 1.2e-9, delta 4.6e-8, gamma 1.1e-7, h 8.8e-13, p 1.5e-7
    bound = {"q": 3e-9, "delta": 1e-7, "gamma": 2.5e-7, "h": 2e-12, "p": 3e-7}
    for k, v in err.items():
        assert v < min(bound[k], 1e-5), (k, v)
