"""The CPU oracle's spectral module against the reference's golden vectors
(src/spectral/test.rs:18-620)."""
import numpy as np
import pytest
from conftest import gold, gold_u64

EXACT = ["hlap", "glap", "rlap", "helm", "c2g2", "simp", "rope", "fope", "filt", "diss", "opak", "rdis",
         "etdv", "htdv", "ap", "etd1", "htd1", "theta", "weight", "hrkx", "hrky", "rk", "spmf"]


@pytest.fixture(scope="module")
def s32(oracle):
    return oracle.Spectral(32, 4)


@pytest.mark.parametrize("name", EXACT)
def test_init_spectral_exact(s32, name):
    # src/spectral/test.rs:325-560: assert_eq!
    assert np.array_equal(s32.get(name), gold(f"spectral/testdata/init_spectral/32_4_{name}.bin"))


@pytest.mark.parametrize("name", ["xtrig", "ytrig", "alk"])
def test_init_spectral_approx(s32, name):
    # src/spectral/test.rs:562-584: abs 1e-13
    np.testing.assert_allclose(s32.get(name), gold(f"spectral/testdata/init_spectral/32_4_{name}.bin"),
                               rtol=0, atol=1e-13)


def test_init_spectral_ints(s32):
    # src/spectral/test.rs:586-620
    assert list(s32.get("xfactors")) == [0, 2, 1, 0, 0]
    assert list(s32.get("yfactors")) == [0, 2, 1, 0, 0]
    assert np.array_equal(s32.get("kmag").astype(np.uint64), gold_u64("spectral/testdata/init_spectral/32_4_kmag.bin"))
    assert s32.get("kmax")[0] == 23 and s32.get("kmaxred")[0] == 16


def test_filter_is_idempotent_and_two_thirds(oracle):
    s = oracle.Spectral(48, 4)
    filt, rk = s.get("filt").reshape(48, 48), s.get("rk")
    rks = rk[None, :] ** 2 + rk[:, None] ** 2
    assert np.array_equal(filt, (rks <= (48 / 3) ** 2).astype(float))


@pytest.mark.parametrize("ng,nz", [(18, 2), (30, 4), (48, 6)])
def test_main_invert(oracle, ng, nz):
    # src/spectral/test.rs:18-120: abs 1e-10
    s = oracle.Spectral(ng, nz)
    p = f"spectral/testdata/main_invert/{ng}_{nz}_"
    qs, ds, gs = gold(p + "qs.bin"), gold(p + "ds.bin"), gold(p + "gs.bin")
    r, u, v, zeta = s.field(), s.field(), s.field(), s.field()
    s.main_invert(qs, ds, gs, r, u, v, zeta)
    for name, a in (("r", r), ("u", u), ("v", v), ("zeta", zeta)):
        np.testing.assert_allclose(a, gold(p + name + ".bin"), rtol=0, atol=1e-10, err_msg=name)


def test_main_invert_linear_case(oracle):
    # delta = gamma = 0 and identical layers: zeta_hat = rope * q_hat (derived from
    # src/spectral/mod.rs:323-375), a closed form independent of the fixtures
    ng, nz = 32, 4
    s = oracle.Spectral(ng, nz)
    rng = np.random.default_rng(3)
    q2 = rng.standard_normal(ng * ng) * s.get("filt")
    qs = np.tile(q2, nz + 1)
    zero = np.zeros_like(qs)
    r, u, v, zeta = s.field(), s.field(), s.field(), s.field()
    s.main_invert(qs, zero, zero.copy(), r, u, v, zeta)
    zs = s.field()
    s.ptospc3d(zeta.copy(), zs, 0, nz)
    np.testing.assert_allclose(zs, np.tile(s.get("rope") * q2, nz + 1), atol=1e-12)


@pytest.mark.parametrize("fn", ["jacob", "divs"])
@pytest.mark.parametrize("ng,nz", [(30, 4), (48, 6)])
def test_jacob_divs(oracle, fn, ng, nz):
    # src/spectral/test.rs:123-180: abs 1e-10
    s = oracle.Spectral(ng, nz)
    p = f"spectral/testdata/{fn}/{ng}_{nz}_"
    cs = s.slab()
    getattr(s, fn)(gold(p + "aa.bin"), gold(p + "bb.bin"), cs)
    np.testing.assert_allclose(cs, gold(p + "cs.bin"), rtol=0, atol=1e-10)


@pytest.mark.parametrize("ng,nz,izend", [(18, 2, 1), (30, 4, 3)])
def test_ptospc3d_spctop3d(oracle, ng, nz, izend):
    # src/spectral/test.rs:183-240: abs 1e-10
    s = oracle.Spectral(ng, nz)
    p = f"spectral/testdata/ptospc3d/{ng}_{nz}_"
    fp, fs = gold(p + "fp.bin"), gold(p + "fs.bin")
    s.ptospc3d(fp, fs, 0, izend)
    np.testing.assert_allclose(fs, gold(p + "fs2.bin"), rtol=0, atol=1e-10)
    np.testing.assert_allclose(fp, gold(p + "fp2.bin"), rtol=0, atol=1e-10)
    p = f"spectral/testdata/spctop3d/{ng}_{nz}_"
    fp, fs = gold(p + "fp.bin"), gold(p + "fs.bin")
    s.spctop3d(fs, fp, 0, izend)
    np.testing.assert_allclose(fs, gold(p + "fs2.bin"), rtol=0, atol=1e-10)
    np.testing.assert_allclose(fp, gold(p + "fp2.bin"), rtol=0, atol=1e-10)


@pytest.mark.parametrize("ng,nz", [(18, 2), (30, 4)])
def test_deal3d(oracle, ng, nz):
    s = oracle.Spectral(ng, nz)
    fp = gold(f"spectral/testdata/deal3d/{ng}_{nz}_fp.bin")
    s.deal3d(fp)
    np.testing.assert_allclose(fp, gold(f"spectral/testdata/deal3d/{ng}_{nz}_fp2.bin"), rtol=0, atol=1e-10)


@pytest.mark.parametrize("ng,nz", [(18, 2), (32, 4)])
def test_deal2d(oracle, ng, nz):
    s = oracle.Spectral(ng, nz)
    fp = gold(f"spectral/testdata/deal2d/{ng}_{nz}_fp.bin")
    s.deal2d(fp)
    np.testing.assert_allclose(fp, gold(f"spectral/testdata/deal2d/{ng}_{nz}_fp2.bin"), rtol=0, atol=1e-10)
    s.deal2d(fp2 := fp.copy())
    np.testing.assert_allclose(fp2, fp, atol=1e-12)  # filtering is idempotent
