"""The CPU oracle's 1-D FFT against the reference's golden vectors (src/stafft/mod.rs:463-876),
plus the closed form (numpy DFT with the Hermitian packing) for sizes the fixtures leave open."""
import json
import os

import numpy as np
import pytest
from conftest import GOLD, gold

FACT = {12: [1, 0, 1, 0, 0], 15: [0, 0, 0, 1, 1], 16: [0, 2, 0, 0, 0], 18: [1, 0, 0, 1, 0], 24: [1, 1, 0, 0, 0]}


def snap(name):
    return json.load(open(os.path.join(GOLD, "snapshots.json")))[name]


def test_factorisen_snapshots(oracle):
    # src/stafft/mod.rs:463-482 (insta snapshots)
    assert oracle.factorisen(16) == (0, snap("shallow_water__stafft__test__factorisen_snapshot_1.snap"))
    assert oracle.factorisen(8640) == (0, snap("shallow_water__stafft__test__factorisen_snapshot_2.snap"))
    assert oracle.factorisen(512)[1] == [0, 4, 1, 0, 0]
    assert oracle.factorisen(1536)[1] == [1, 4, 0, 0, 0]
    assert oracle.factorisen(14)[0] != 0  # prime factor 7: the reference exits(1)


@pytest.mark.parametrize("n", [18, 30, 32])
def test_initfft(oracle, n):
    # src/stafft/mod.rs:484-527, abs 1e-13
    f, trig = oracle.initfft(n)
    np.testing.assert_allclose(trig, gold(f"stafft/testdata/initfft/{n}_trig.bin"), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [12, 15, 16, 18, 24])
@pytest.mark.parametrize("case", [1, 2])
def test_forfft_exact(oracle, n, case):
    # src/stafft/mod.rs:529-697: assert_eq! (bit-exact)
    p = f"stafft/testdata/forfft/forfft_ng{n}_{case}_"
    x, x2, trig = gold(p + "x.bin"), gold(p + "x2.bin"), gold(p + "trig.bin")
    assert oracle.forfft(n, n, x, trig, FACT[n]) == 0
    assert np.array_equal(x, x2, equal_nan=True)


@pytest.mark.parametrize("n", [12, 15, 16, 18, 24])
@pytest.mark.parametrize("case", [1, 2])
def test_revfft_exact(oracle, n, case):
    # src/stafft/mod.rs:699-876: assert_eq! (bit-exact)
    p = f"stafft/testdata/revfft/revfft_ng{n}_{case}_"
    x, x2, trig = gold(p + "x.bin"), gold(p + "x2.bin"), gold(p + "trig.bin")
    assert oracle.revfft(n, n, x, trig, FACT[n]) == 0
    assert np.array_equal(x, x2, equal_nan=True)


def hermitian_pack(F, n):
    out = np.zeros(F.shape)
    for k in range(n // 2 + 1):
        out[k] = F[k].real
    for k in range(1, (n + 1) // 2):
        out[n - k] = F[k].imag
    return out


@pytest.mark.parametrize("n", [6, 8, 9, 12, 20, 32, 36, 45, 48, 64, 96, 120, 128, 256, 384, 512, 1536])
def test_forfft_closed_form(oracle, n):
    # out[k] = Re F_k / sqrt(n), out[n-k] = Im F_k / sqrt(n), F_k = sum_j x_j exp(-2 pi i j k / n)
    # batch m = n as in every solver call: with m <= (lv-1)/2 the reference takes its small-batch
    # branches, whose radix-6 twiddle indexing is not a DFT (forward.rs:103-106, reverse.rs:143-148)
    rng = np.random.default_rng(n)
    m = n
    x = rng.standard_normal((n, m))
    f, trig = oracle.initfft(n)
    xs = x.ravel().copy()
    assert oracle.forfft(m, n, xs, trig, f) == 0
    want = hermitian_pack(np.fft.fft(x, axis=0) / np.sqrt(n), n)
    np.testing.assert_allclose(xs.reshape(n, m), want, rtol=0, atol=1e-13 * np.sqrt(n))
    # the inverse restores the input
    assert oracle.revfft(m, n, xs, trig, f) == 0
    np.testing.assert_allclose(xs.reshape(n, m), x, rtol=0, atol=1e-13 * np.sqrt(n))


def test_radix2_after_radix5_is_rejected(oracle):
    # n = 10 factorises as [0,0,1,0,1]; the reference's forrdx2 indexes cosine[[0,k-1]] and panics
    f, trig = oracle.initfft(10)
    x = np.ones(10)
    assert oracle.forfft(1, 10, x, trig, f) != 0
