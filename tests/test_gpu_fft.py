"""GPU parity of the 1-D and 2-D transforms and spectral derivatives (stafft, sta2dfft) against
the reference's golden vectors, the oracle and the closed form.  The GPU transform is a different
algorithm from the reference's (Stockham in LDS vs Temperton stages), so results agree to rounding:
tolerances are relative to the size of the data, 1e-14 scaled by log2(n)."""
import numpy as np
import pytest
from conftest import gold

pytestmark = pytest.mark.gpu
FACT = {12: [1, 0, 1, 0, 0], 15: [0, 0, 0, 1, 1], 16: [0, 2, 0, 0, 0], 18: [1, 0, 0, 1, 0], 24: [1, 1, 0, 0, 0]}


def close(a, b, tol):
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b)) / scale
    assert err <= tol, f"max rel err {err:.3e} > {tol:.1e}"


@pytest.mark.parametrize("n", [12, 15, 16, 18, 24])
@pytest.mark.parametrize("case", [1, 2])
@pytest.mark.parametrize("kind", ["forfft", "revfft"])
def test_fft_golden(sw, n, case, kind):
    # src/stafft/mod.rs:529-876 (bit-exact there; rounding-level here)
    p = f"stafft/testdata/{kind}/{kind}_ng{n}_{case}_"
    x, want = gold(p + "x.bin"), gold(p + "x2.bin")
    getattr(sw, kind)(n, n, x, gold(p + "trig.bin"), FACT[n])
    close(x, want, 1e-14)


def pack(F, n):
    out = np.zeros(F.shape)
    for k in range(n // 2 + 1):
        out[k] = F[k].real
    for k in range(1, (n + 1) // 2):
        out[n - k] = F[k].imag
    return out


@pytest.mark.parametrize("n,m", [(16, 16), (32, 7), (64, 48), (96, 96), (120, 30), (128, 128), (256, 33),
                                 (384, 16), (512, 512), (1536, 64), (30, 30), (48, 1), (45, 9)])
def test_forfft_closed_form_and_roundtrip(sw, n, m):
    rng = np.random.default_rng(n + m)
    x = rng.standard_normal((n, m))
    xs = x.ravel().copy()
    sw.forfft(m, n, xs)
    close(xs.reshape(n, m), pack(np.fft.fft(x, axis=0) / np.sqrt(n), n), 2e-15 * np.log2(n))
    sw.revfft(m, n, xs)
    close(xs.reshape(n, m), x, 4e-15 * np.log2(n))


def test_trig_must_be_initffts_table(sw):
    """forfft/revfft use the twiddles of initfft(n) (stafft/mod.rs:67-95): a table of another length or
    content is refused, from host or device memory (device arrays allocated through the HIP runtime the
    library is linked against), and the array is left untouched."""
    import ctypes as C
    n, m = 16, 4
    _, trig = sw.initfft(n)
    x = np.arange(n * m, dtype=np.float64)
    sw.forfft(m, n, x.copy(), trig)
    for bad in (sw.initfft(32)[1][:2 * n].copy(), trig * (1 + 1e-9)):
        y = x.copy()
        with pytest.raises(sw.SwError) as e:
            sw.forfft(m, n, y, bad)
        assert e.value.code == -1 and "initfft" in str(e.value)
        assert np.array_equal(y, x)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    fwd = sw._lib.sw_forfft
    fwd.argtypes = [C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
    ptrs = []
    try:
        for a in (x, trig, trig + np.where(np.arange(2 * n) == 3, 1.0, 0.0)):
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
            ptrs.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        xd, td, bad = ptrs
        assert fwd(m, n, xd, td, None) == 0
        assert fwd(m, n, xd, bad, None) == -1
    finally:
        for p in ptrs:
            hip.hipFree(p)


def test_unsupported_lengths(sw):
    x = np.zeros(10 * 10)
    with pytest.raises(sw.SwError) as e:
        sw.forfft(10, 10, x)  # radix 2 after radix 5: the reference panics
    assert e.value.code == -5
    with pytest.raises(sw.SwError) as e:
        sw.forfft(2, 14, np.zeros(28))
    assert e.value.code == -4


def test_ptospc_spctop_golden(sw):
    # src/sta2dfft.rs:252-316
    d = sw.D2FFT(32, 32)
    rvar, svar = gold("testdata/ptospc/ptospc_ng32_rvar.bin"), np.zeros(32 * 32)
    d.ptospc(rvar, svar)
    close(svar, gold("testdata/ptospc/ptospc_ng32_svar2.bin"), 1e-14)
    svar, rvar = gold("testdata/spctop/spctop_ng32_svar.bin"), np.zeros(32 * 32)
    d.spctop(svar, rvar)
    close(rvar, gold("testdata/spctop/spctop_ng32_rvar2.bin"), 1e-14)


@pytest.mark.parametrize("ng", [18, 24, 30, 48, 64, 96, 128, 256, 512, 1536])
def test_2d_against_oracle(sw, oracle, ng):
    rng = np.random.default_rng(ng)
    phys = rng.standard_normal(ng * ng)
    xf, yf, xt, yt, kx, ky = oracle.init2dfft(ng, ng, 2 * np.pi, 2 * np.pi)
    want = np.zeros(ng * ng)
    oracle.ptospc(ng, ng, xf, yf, xt, yt, phys.copy(), want)
    d = sw.D2FFT(ng, ng)
    got = np.zeros(ng * ng)
    d.ptospc(phys, got)
    close(got, want, 2e-15 * np.log2(ng) * 2)
    back = np.zeros(ng * ng)
    d.spctop(got, back)
    close(back, phys, 4e-15 * np.log2(ng) * 2)


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("case,n", [(1, 32), (2, 24)])
def test_deriv_golden_exact(sw, axis, case, n):
    # src/sta2dfft.rs:318-390: assert_eq! (pure index permutation and multiply: exact here too)
    p = f"testdata/deriv/{axis}deriv_{case}_"
    d = sw.D2FFT(n, n)
    der = np.zeros(n * n)
    getattr(d, f"{axis}deriv")(gold(p + f"rk{axis}.bin"), gold(p + "var.bin"), der)
    assert np.array_equal(der, gold(p + "der2.bin"))
