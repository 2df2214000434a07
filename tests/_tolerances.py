"""Parity bars of the nhswps step, shared by the GPU tests and the smoke check.

Per-field relative L2 bound of the GPU result against the oracle after psolve or one advance step, for
the reference's golden states and synthetic states up to ng = 128 (larger grids are held to a bound
measured from the oracle's own sensitivity in the same test: tests/test_gpu_shapes.py).  1e-12 where
the step is well conditioned.  The pressure solve amplifies rounding differences: the reference
algorithm itself moves pn by ~1.5e-11, ds by ~2.4e-12 and w by ~2e-12 after one step of its 24x4
golden state when its inputs are perturbed by 1e-16 (tests/test_oracle_conditioning.py asserts these
figures), and a transform that rounds differently perturbs every one of the step's thousands of
transforms at that level.  The bounds are about ten times the largest such figure, or the largest
GPU error measured on MI355X, whichever is larger.
"""
STEP_TOL = {"qs": 1e-12, "r": 1e-12, "u": 1e-12, "v": 1e-12, "zeta": 1e-12, "z": 1e-12, "ri": 1e-12,
            "zx": 5e-12, "zy": 5e-12, "aa": 5e-12, "w": 5e-11, "ds": 1e-10, "gs": 5e-11,
            "pn": 5e-10, "ps": 5e-10, "dpn": 5e-10}
