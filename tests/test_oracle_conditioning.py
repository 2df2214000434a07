"""How strongly one nhswps step amplifies rounding-level differences, measured on the oracle alone.

This is what fixes the parity tolerances of the GPU step tests: an implementation whose transforms
round differently from the reference's (by ~1e-16 relative, as any other FFT does) cannot agree
with it more closely than the step's own sensitivity to such perturbations.  The test pins that
sensitivity from both sides: perturbing qs, ds, gs of the reference's 24x4 advance state by 1e-16
relative noise moves pn by ~1.5e-11, ds by ~2.4e-12 and w by ~2e-12 after one step."""
import numpy as np
from conftest import gold, rel_l2

ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]


def test_one_step_sensitivity_to_rounding_noise(oracle):
    ng, nz, t0 = 24, 4, 0.624_999_999_999_999_9
    p = f"nhswps/testdata/advance/{ng}_{nz}/"
    base = {f: gold(p + f + ".bin") for f in ALL}
    s = oracle.Spectral(ng, nz)
    a = oracle.State(s, t=t0, ggen=True, **base)
    oracle.advance(a)
    rng = np.random.default_rng(0)
    pert = {f: (v * (1 + 1e-16 * rng.standard_normal(v.shape)) if f in ("qs", "ds", "gs") else v)
            for f, v in base.items()}
    b = oracle.State(s, t=t0, ggen=True, **pert)
    rc, _, iters = oracle.advance(b)
    assert rc == 0
    d = {f: rel_l2(getattr(b, f), getattr(a, f)) for f in ("qs", "ds", "pn", "w")}
    # q stays at the noise level; the pressure is amplified by ~1e5.  Measured (seeds 0-3): pn 1.1e-11 to
    # 1.5e-11, ds 1.1e-12 to 2.5e-12, w 1.0e-12 to 2.1e-12; asserted within a factor of about three.
    assert d["qs"] < 1e-15
    assert 5e-12 < d["pn"] < 5e-11, d
    assert 5e-13 < d["ds"] < 8e-12, d
    assert 4e-13 < d["w"] < 7e-12, d
