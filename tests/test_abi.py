"""CPU checks of the drop-in boundary: the HIP library loads, exports every entry point declared
in include/shallow_water.h, and its host-only entry points (factorisation, trig table, synthetic
generator) agree with the oracle.  No GPU work is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest
from conftest import ROOT

HDR = os.path.join(ROOT, "include", "shallow_water.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char\*)\s+(sw_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_reference_surface():
    names = declared()
    for n in ("sw_forfft", "sw_revfft", "sw_ptospc", "sw_spctop", "sw_xderiv", "sw_yderiv", "sw_main_invert",
              "sw_deal2d", "sw_deal3d", "sw_advance", "sw_source", "sw_psolve", "sw_coeffs", "sw_vertical",
              "sw_ensemble_create", "sw_ensemble_advance", "sw_ensemble_destroy", "sw_balinit", "sw_balinit_file"):
        assert n in names


def test_library_exports_every_declared_symbol(sw):
    lib = C.CDLL(sw.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("n", [16, 18, 30, 32, 120, 512, 1536, 8640])
def test_factorisen_and_initfft_match_oracle(sw, oracle, n):
    assert sw.factorisen(n) == oracle.factorisen(n)[1]
    f, trig = sw.initfft(n)
    f2, trig2 = oracle.initfft(n)
    assert f == [int(x) for x in f2]
    assert np.array_equal(trig, trig2)


def test_factorisen_rejects_prime_7(sw):
    with pytest.raises(sw.SwError) as e:
        sw.factorisen(14)
    assert e.value.code == -4


def test_synthetic_field_is_seeded_and_scaled(sw, oracle):
    ng, nz = 64, 4
    qs, ds = sw.synth_kminus3(ng, nz, seed=1, k0=8.0, zeta_rms=2 * np.pi)
    qs2, _ = sw.synth_kminus3(ng, nz, seed=1, k0=8.0, zeta_rms=2 * np.pi)
    assert np.array_equal(qs, qs2) and not ds.any()
    s = oracle.Spectral(ng, nz)
    # identical levels, de-aliased, zero mean
    q = qs.reshape(nz + 1, ng * ng)
    assert all(np.array_equal(q[0], q[i]) for i in range(nz + 1))
    assert np.all(q[0][s.get("filt") == 0] == 0) and q[0][0] == 0
    # the oracle's inversion gives rms(zeta) = 2 pi
    r, u, v, zeta = s.field(), s.field(), s.field(), s.field()
    s.main_invert(qs, ds, ds.copy(), r, u, v, zeta)
    assert abs(np.sqrt(np.mean(zeta ** 2)) - 2 * np.pi) < 1e-9


def test_read_r8_reads_values_after_the_header_and_checks_the_size(sw, tmp_path):
    vals = np.arange(12, dtype="<f8") * 0.25 - 1.0
    p = tmp_path / "x_init.r8"
    p.write_bytes(np.zeros(1, "<f8").tobytes() + vals.tobytes())
    assert np.array_equal(sw.read_r8(str(p), 12), vals)
    with pytest.raises(sw.SwError) as e:
        sw.read_r8(str(p), 13)
    assert e.value.code == -1 and "expected an 8-byte header" in str(e.value)
    with pytest.raises(sw.SwError):
        sw.read_r8(str(tmp_path / "missing.r8"), 1)


def test_fft_tile_override_is_validated(sw):
    # host-only: the tile table and its overrides (no device work)
    rule = sw.fft_get_tile(96, 8)
    assert rule[0] >= 1 and rule[1] % 64 == 0
    for bad in ((5, 64), (1, 128), (4, 96), (4, 576), (6, 64 * 9)):
        with pytest.raises(sw.SwError):
            sw.fft_set_tile(96, 8, *bad)
    try:
        sw.fft_set_tile(96, 8, 2, 128)
        assert sw.fft_get_tile(96, 8) == (2, 128)
        assert sw.fft_get_tile(96, 4) != (2, 128) or rule == (2, 128)
    finally:
        sw.fft_set_tile(96, 8, 0, 0)
    assert sw.fft_get_tile(96, 8) == rule
