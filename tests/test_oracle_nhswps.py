"""The CPU oracle's nhswps kernels against the reference's golden vectors
(src/nhswps/{vertical,coeffs,cpsource,source,psolve,advance}.rs test modules)."""
import numpy as np
import pytest
from conftest import gold

D = "nhswps/testdata/"


def state(oracle, ng, nz, prefix, names, **kw):
    s = oracle.Spectral(ng, nz)
    return oracle.State(s, **{n: gold(prefix + n + ".bin") for n in names}, **kw)


@pytest.mark.parametrize("ng,nz", [(18, 2), (32, 4)])
def test_vertical(oracle, ng, nz):
    # src/nhswps/vertical.rs:237-308: abs 1e-10
    p = f"{D}vertical/{ng}_{nz}_"
    st = state(oracle, ng, nz, p, ["z", "zx", "zy", "r", "w", "aa", "u", "v", "ds"])
    oracle.vertical(st)
    for f in ("z", "zx", "zy", "w", "aa"):
        np.testing.assert_allclose(getattr(st, f), gold(p + f + "2.bin"), rtol=0, atol=1e-10, err_msg=f)


@pytest.mark.parametrize("ng,nz", [(18, 2), (32, 4)])
def test_coeffs(oracle, ng, nz):
    # src/nhswps/coeffs.rs:184-333: abs 1e-10
    p = f"{D}coeffs/{ng}_{nz}_"
    st = state(oracle, ng, nz, p, ["ri", "zx", "zy"])
    out = {n: gold(p + n + ".bin") for n in ("sigx", "sigy", "cpt1", "cpt2")}
    oracle.coeffs(st, out["sigx"], out["sigy"], out["cpt1"], out["cpt2"])
    for n, a in out.items():
        np.testing.assert_allclose(a, gold(p + n + "2.bin"), rtol=0, atol=1e-10, err_msg=n)


@pytest.mark.parametrize("ng,nz", [(18, 2), (32, 4)])
def test_cpsource(oracle, ng, nz):
    # src/nhswps/cpsource.rs:351-381: abs 1e-13
    p = f"{D}cpsource/{ng}_{nz}_"
    st = state(oracle, ng, nz, p, ["ri", "u", "v", "w", "zeta", "z", "zx", "zy"])
    sp0 = gold(p + "sp0.bin")
    oracle.cpsource(st, sp0)
    np.testing.assert_allclose(sp0, gold(p + "sp02.bin"), rtol=0, atol=1e-13)


@pytest.mark.parametrize("ng,nz", [(18, 2), (32, 4)])
def test_source(oracle, ng, nz):
    # src/nhswps/source.rs:236-406: abs 1e-13
    p = f"{D}source/{ng}_{nz}_"
    st = state(oracle, ng, nz, p, ["aa", "qs", "ds", "ps", "u", "v", "ri", "dpn", "z", "zx", "zy"])
    out = {n: gold(p + n + ".bin") for n in ("sqs", "sds", "sgs")}
    oracle.source(st, out["sqs"], out["sds"], out["sgs"])
    for n, a in out.items():
        np.testing.assert_allclose(a, gold(p + n + "2.bin"), rtol=0, atol=1e-13, err_msg=n)


ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]


@pytest.mark.parametrize("ng,nz", [(24, 4), (32, 4)])
def test_psolve(oracle, ng, nz):
    # src/nhswps/psolve.rs:524-625: abs 1e-10
    p = f"{D}psolve/{ng}_{nz}_"
    st = state(oracle, ng, nz, p, ALL)
    rc, iters, errp = oracle.psolve(st)
    assert rc == 0 and 1 <= iters < 100 and errp <= 1e-9
    for f in ("z", "zx", "zy", "w", "aa", "ri", "pn", "ps", "dpn"):
        np.testing.assert_allclose(getattr(st, f), gold(p + f + "2.bin"), rtol=0, atol=1e-10, err_msg=f)


def test_psolve_iteration_cap(oracle):
    # the reference exits(1) when the cap is reached (psolve.rs:360-366); here an error code
    p = f"{D}psolve/24_4_"
    st = state(oracle, 24, 4, p, ALL)
    rc, iters, errp = oracle.psolve(st, nitmax=1)
    assert rc == -2 and iters == 1


@pytest.mark.parametrize("ng,nz,t0,t1", [(18, 6, 5.555_555_555_555_555E-2, 0.111_111_111_111_111_1),
                                         (24, 4, 0.624_999_999_999_999_9, 0.624_999_999_999_999_9 + 1 / 24)])
def test_advance(oracle, ng, nz, t0, t1):
    # src/nhswps/advance.rs:296-600: 16 fields and t at abs 1e-10
    p = f"{D}advance/{ng}_{nz}/"
    st = state(oracle, ng, nz, p, ALL, t=t0, ggen=True)
    rc, diag, iters = oracle.advance(st)
    assert rc == 0 and all(i >= 1 for i in iters)
    assert abs(st.t - t1) < 1e-10
    for f in ALL:
        np.testing.assert_allclose(getattr(st, f), gold(p + f + "2.bin"), rtol=0, atol=1e-10, err_msg=f)
