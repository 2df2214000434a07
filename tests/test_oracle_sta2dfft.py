"""The CPU oracle's 2-D transform and spectral derivatives against the reference's golden vectors
(src/sta2dfft.rs:196-390)."""
import json
import math
import os

import numpy as np
import pytest
from conftest import GOLD, gold


def snap(name):
    return json.load(open(os.path.join(GOLD, "snapshots.json")))[name]


@pytest.mark.parametrize("ng", [30, 120])
def test_init2dfft(oracle, ng):
    # src/sta2dfft.rs:198-250: trig within f64::EPSILON, factors and wavenumbers as snapshots
    lx = 6.283_185_307_179_586
    xf, yf, xt, yt, kx, ky = oracle.init2dfft(ng, ng, lx, lx)
    want = gold(f"testdata/init2dfft/ng{ng}_trig.bin")
    np.testing.assert_allclose(xt, want, rtol=0, atol=np.finfo(float).eps)
    np.testing.assert_allclose(yt, want, rtol=0, atol=np.finfo(float).eps)
    base = f"shallow_water__sta2dfft__test__init2dfft_ng{ng}"
    assert list(xf) == snap(base + ".snap")
    assert list(yf) == snap(base + "-2.snap")
    assert list(kx) == snap(base + "-3.snap")
    assert list(ky) == snap(base + "-4.snap")


def test_ptospc_ng32_exact(oracle):
    # src/sta2dfft.rs:252-283: assert_eq! on both the output and the destroyed input
    p = "testdata/ptospc/ptospc_ng32_"
    rvar, svar, trig = gold(p + "rvar.bin"), gold(p + "svar.bin"), gold(p + "trig.bin")
    f = [0, 2, 1, 0, 0]
    oracle.ptospc(32, 32, f, f, trig, trig, rvar, svar)
    assert np.array_equal(rvar, gold(p + "rvar2.bin"))
    assert np.array_equal(svar, gold(p + "svar2.bin"))


def test_spctop_ng32_exact(oracle):
    # src/sta2dfft.rs:285-316
    p = "testdata/spctop/spctop_ng32_"
    rvar, svar, trig = gold(p + "rvar.bin"), gold(p + "svar.bin"), gold(p + "trig.bin")
    f = [0, 2, 1, 0, 0]
    oracle.spctop(32, 32, f, f, trig, trig, svar, rvar)
    assert np.array_equal(rvar, gold(p + "rvar2.bin"))
    assert np.array_equal(svar, gold(p + "svar2.bin"))


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("case,n", [(1, 32), (2, 24)])
def test_deriv_exact(oracle, axis, case, n):
    # src/sta2dfft.rs:318-390: assert_eq!
    p = f"testdata/deriv/{axis}deriv_{case}_"
    rk = gold(p + f"rk{axis}.bin")
    var, der = gold(p + "var.bin"), gold(p + "der.bin")
    (oracle.xderiv if axis == "x" else oracle.yderiv)(n, n, rk, var, der)
    assert np.array_equal(der, gold(p + "der2.bin"))


def test_derivative_of_cosine(oracle):
    # d/dx cos(3x) = -3 sin(3x): a closed-form pin of the derivative sign and wavenumber
    ng = 32
    xf, yf, xt, yt, kx, ky = oracle.init2dfft(ng, ng, 2 * math.pi, 2 * math.pi)
    x = -math.pi + 2 * math.pi * np.arange(ng) / ng
    X, Y = np.meshgrid(x, x, indexing="ij")  # X[j, i] -> element (i=y, j=x) at i + ng*j
    phys = np.cos(3 * X) * np.cos(2 * Y)
    rvar = phys.ravel().copy()
    svar = np.zeros(ng * ng)
    oracle.ptospc(ng, ng, xf, yf, xt, yt, rvar, svar)
    der = np.zeros(ng * ng)
    oracle.xderiv(ng, ng, kx, svar, der)
    out = np.zeros(ng * ng)
    oracle.spctop(ng, ng, xf, yf, xt, yt, der, out)
    np.testing.assert_allclose(out.reshape(ng, ng), -3 * np.sin(3 * X) * np.cos(2 * Y), atol=1e-12)
    oracle.ptospc(ng, ng, xf, yf, xt, yt, phys.ravel().copy(), svar)
    oracle.yderiv(ng, ng, ky, svar, der)
    oracle.spctop(ng, ng, xf, yf, xt, yt, der, out)
    np.testing.assert_allclose(out.reshape(ng, ng), -2 * np.cos(3 * X) * np.sin(2 * Y), atol=1e-12)
