"""GPU parity of the spectral module (src/spectral/mod.rs) against the reference's golden vectors
(their own tolerances) and against the oracle (relative L2, rounding level)."""
import numpy as np
import pytest
from conftest import gold, rel_l2

pytestmark = pytest.mark.gpu

EXACT = ["hlap", "glap", "rlap", "helm", "c2g2", "simp", "rope", "fope", "filt", "diss", "opak", "rdis",
         "etdv", "htdv", "ap", "etd1", "htd1", "theta", "weight", "hrkx", "hrky", "rk", "spmf"]


def test_operators_exact(sw):
    # src/spectral/test.rs:325-560: the library builds Spectral::new's operators itself
    s = sw.Spectral(32, 4)
    for name in EXACT:
        assert np.array_equal(s.get(name), gold(f"spectral/testdata/init_spectral/32_4_{name}.bin")), name


@pytest.mark.parametrize("ng,nz", [(18, 2), (30, 4), (48, 6)])
def test_main_invert(sw, oracle, ng, nz):
    s = sw.Spectral(ng, nz)
    o = oracle.Spectral(ng, nz)
    p = f"spectral/testdata/main_invert/{ng}_{nz}_"
    qs, ds, gs = gold(p + "qs.bin"), gold(p + "ds.bin"), gold(p + "gs.bin")
    out = [s.field() for _ in range(4)]
    ref = [o.field() for _ in range(4)]
    s.main_invert(qs, ds, gs, *out)
    o.main_invert(qs, ds, gs, *ref)
    for name, a, b in zip(("r", "u", "v", "zeta"), out, ref):
        np.testing.assert_allclose(a, gold(p + name + ".bin"), rtol=0, atol=1e-10, err_msg=name)
        assert rel_l2(a, b) < 1e-12, name


@pytest.mark.parametrize("fn", ["jacob", "divs"])
@pytest.mark.parametrize("ng,nz", [(30, 4), (48, 6)])
def test_jacob_divs(sw, oracle, fn, ng, nz):
    s = sw.Spectral(ng, nz)
    p = f"spectral/testdata/{fn}/{ng}_{nz}_"
    cs = s.slab()
    getattr(s, fn)(gold(p + "aa.bin"), gold(p + "bb.bin"), cs)
    np.testing.assert_allclose(cs, gold(p + "cs.bin"), rtol=0, atol=1e-10)
    want = s.slab()
    getattr(oracle.Spectral(ng, nz), fn)(gold(p + "aa.bin"), gold(p + "bb.bin"), want)
    assert rel_l2(cs, want) < 1e-12


@pytest.mark.parametrize("ng,nz,izend", [(18, 2, 1), (30, 4, 3)])
def test_ptospc3d_spctop3d(sw, ng, nz, izend):
    s = sw.Spectral(ng, nz)
    p = f"spectral/testdata/ptospc3d/{ng}_{nz}_"
    fs = gold(p + "fs.bin")
    s.ptospc3d(gold(p + "fp.bin"), fs, 0, izend)
    np.testing.assert_allclose(fs, gold(p + "fs2.bin"), rtol=0, atol=1e-10)
    p = f"spectral/testdata/spctop3d/{ng}_{nz}_"
    fp = gold(p + "fp.bin")
    s.spctop3d(gold(p + "fs.bin"), fp, 0, izend)
    np.testing.assert_allclose(fp, gold(p + "fp2.bin"), rtol=0, atol=1e-10)


@pytest.mark.parametrize("ng,nz", [(18, 2), (30, 4)])
def test_deal3d(sw, ng, nz):
    s = sw.Spectral(ng, nz)
    fp = gold(f"spectral/testdata/deal3d/{ng}_{nz}_fp.bin")
    s.deal3d(fp)
    np.testing.assert_allclose(fp, gold(f"spectral/testdata/deal3d/{ng}_{nz}_fp2.bin"), rtol=0, atol=1e-10)


@pytest.mark.parametrize("ng,nz", [(18, 2), (32, 4)])
def test_deal2d(sw, ng, nz):
    s = sw.Spectral(ng, nz)
    fp = gold(f"spectral/testdata/deal2d/{ng}_{nz}_fp.bin")
    s.deal2d(fp)
    np.testing.assert_allclose(fp, gold(f"spectral/testdata/deal2d/{ng}_{nz}_fp2.bin"), rtol=0, atol=1e-10)


def test_main_invert_large_against_oracle(sw, oracle):
    ng, nz = 128, 8
    qs, ds = sw.synth_kminus3(ng, nz, seed=5, k0=8.0, delta_rms=0.1)
    gs = np.zeros_like(qs)
    s, o = sw.Spectral(ng, nz), oracle.Spectral(ng, nz)
    out = [s.field() for _ in range(4)]
    ref = [o.field() for _ in range(4)]
    s.main_invert(qs, ds, gs, *out)
    o.main_invert(qs, ds, gs, *ref)
    for a, b in zip(out, ref):
        assert rel_l2(a, b) < 1e-12
