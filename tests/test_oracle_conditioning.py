"""How strongly one nhswps step amplifies rounding-level differences, measured on the oracle alone.

This is what fixes the parity tolerances of the GPU step tests: an implementation whose transforms
round differently from the reference's (by ~1e-16 relative, as any other FFT does) cannot agree
with it more closely than the step's own sensitivity to such perturbations."""
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
    # q stays at the noise level; the pressure is amplified by ~1e5
    assert d["qs"] < 1e-15
    assert d["pn"] > 1e-12 and d["ds"] > 1e-13 and d["w"] > 1e-13, d
