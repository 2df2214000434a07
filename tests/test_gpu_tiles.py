"""The tile of the block-per-rows transform passes (rows per block, threads per block; sw_fft_set_tile,
chosen per length by autotune.py) changes only the speed: transforms and a whole step are bit-identical
across valid tiles, in both precisions.  Also the tuner's timing entry point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]


def tiles(ng, nbytes):
    out = []
    for b in (2, 4, 6, 8, 16, 32, ng):
        for nt in (64, 128, 256, 512):
            if ng % b or b % 2 or max(b, 2) * ng * nbytes > 163840:
                continue
            if (b // 2) * ng <= 16 * nt:
                out.append((b, nt))
    return sorted(set(out))


def step_fields(sw, ng, nz, precision):
    qs, ds = sw.synth_kminus3(ng, nz, seed=3, k0=8.0, zeta_rms=0.3 * 4 * np.pi, delta_rms=0.01 * 4 * np.pi)
    st = sw.State(sw.Spectral(ng, nz, precision), ggen=True, qs=qs, ds=ds, gs=np.zeros_like(qs))
    sw.advance(st)
    return {f: st.get(f) for f in ALL}


@pytest.mark.parametrize("ng,nz", [(96, 4), (64, 8)])
def test_transforms_and_step_do_not_depend_on_tile(sw, ng, nz):
    rng = np.random.default_rng(ng)
    phys = rng.standard_normal(ng * ng * (nz + 1))

    def transforms():
        s = sw.Spectral(ng, nz)
        spec, back = s.field(), s.field()
        s.ptospc3d(phys, spec, 0, nz)
        s.spctop3d(spec, back, 0, nz)
        return spec, back

    base = transforms()
    ref = {p: step_fields(sw, ng, nz, p) for p in (sw.FP64, sw.MIXED)}
    try:
        for nbytes, prec in ((8, sw.FP64), (4, sw.MIXED)):
            cands = tiles(ng, nbytes)
            picked = [cands[0], cands[len(cands) // 2], cands[-1]]
            for b, nt in picked:
                sw.fft_set_tile(ng, nbytes, b, nt)
                assert sw.fft_get_tile(ng, nbytes) == (b, nt)
                if nbytes == 8:
                    spec, back = transforms()
                    assert np.array_equal(spec, base[0]) and np.array_equal(back, base[1]), (b, nt)
                got = step_fields(sw, ng, nz, prec)
                for f in ALL:
                    assert np.array_equal(got[f], ref[prec][f]), (nbytes, b, nt, f)
            sw.fft_set_tile(ng, nbytes, 0, 0)
    finally:
        sw.fft_set_tile(ng, 8, 0, 0)
        sw.fft_set_tile(ng, 4, 0, 0)


def test_time_tile(sw):
    us = sw.fft_time_tile(96, 8, 4, 64, 8, 3)
    assert 0.0 < us < 1e6
    with pytest.raises(sw.SwError):
        sw.fft_time_tile(96, 8, 5, 64, 8, 3)  # B must divide ng and be even
