"""The whole program on the GPU against the reference's end-to-end case (src/test.rs:27-36, 86-260):
sw_nhswps reads qq/dd/gg_init.r8 made from the reference's balanced 32x4 state (testdata/balinit/
32_4_sw_init.r8, each record copied to the nz + 1 levels as swto3d does, src/swto3d.rs:10-39), runs to
t = 1.0 with saves every 0.25 (time step 1/32) and writes the reference's files.  monitor.asc and
ecomp.asc must equal the reference's token for token, as its own test compares them, and the .r4 dumps
must be within the reference's 1e-8 (src/utils.rs:62-79).  spectra.asc, which the reference's tests do
not compare (too noisy for bitwise checks), is compared with the oracle's savegrid spectra.

Also: savegrid's output (sw_savegrid) against the oracle's at a wave-kernel size, and the FP64
checkpoint (sw_state_save / sw_state_load) round trip."""
import os
import shutil

import numpy as np
import pytest
from conftest import GOLD, gold, rel_l2

pytestmark = pytest.mark.gpu
CASE = os.path.join(GOLD, "testdata")
ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]


def write_init_files(out, ng, nz):
    """swto3d: each record of sw_init.r8 (an 8-byte header, then ng^2 f64) copied to nz + 1 levels behind an
    8-byte header."""
    raw = np.fromfile(os.path.join(CASE, "balinit", "32_4_sw_init.r8"), dtype="<f8").reshape(3, ng * ng + 1)
    for name, rec in zip(("qq", "dd", "gg"), raw[:, 1:]):
        with open(os.path.join(out, f"{name}_init.r8"), "wb") as f:
            f.write(np.zeros(1, "<f8").tobytes() + np.tile(rec, nz + 1).astype("<f8").tobytes())


def read_r4(path, n):
    a = np.fromfile(path, dtype="<f4")
    assert a.size % (n + 1) == 0, path
    a = a.reshape(-1, n + 1)
    return a[:, 0], a[:, 1:]


@pytest.fixture(scope="module")
def run_32_4(sw, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("nhswps_32_4"))
    write_init_files(out, 32, 4)
    its = sw.nhswps(out, 32, 4, 1.0 / 32, 1.0, 0.25)
    print(f"nhswps 32x4 to t = 1: {its} psolve iterations", flush=True)
    return out


def test_pipeline_from_qq_init_32_4(sw, tmp_path):
    """The reference's pipeline after vstrip, all on the GPU: balinit from its 32_4_qq_init.r8 (sw_balinit_file
    writes sw_init.r8), the layer copy of swto3d, then nhswps to t = 1.0: monitor.asc and ecomp.asc token
    for token as the reference's."""
    ng, nz = 32, 4
    out = str(tmp_path)
    shutil.copyfile(os.path.join(CASE, "balinit", "32_4_qq_init.r8"), os.path.join(out, "qq_init.r8"))
    assert sw.balinit_file(out, ng, nz) == 9
    raw = np.fromfile(os.path.join(out, "sw_init.r8"), dtype="<f8").reshape(3, ng * ng + 1)
    for name, rec in zip(("qq", "dd", "gg"), raw[:, 1:]):
        with open(os.path.join(out, f"{name}_init.r8"), "wb") as f:
            f.write(np.zeros(1, "<f8").tobytes() + np.tile(rec, nz + 1).astype("<f8").tobytes())
    sw.nhswps(out, ng, nz, 1.0 / 32, 1.0, 0.25)
    for name in ("monitor.asc", "ecomp.asc"):
        mine = open(os.path.join(out, name)).read().split()
        ref = open(os.path.join(CASE, "complete", "32_4", name)).read().split()
        diff = [(i, a, b) for i, (a, b) in enumerate(zip(mine, ref)) if a != b]
        assert len(mine) == len(ref) and not diff, (name, diff[:5])


def test_nhswps_32_4_monitor_and_ecomp(run_32_4):
    for name in ("monitor.asc", "ecomp.asc"):
        mine = open(os.path.join(run_32_4, name)).read().split()
        ref = open(os.path.join(CASE, "complete", "32_4", name)).read().split()
        diff = [(i, a, b) for i, (a, b) in enumerate(zip(mine, ref)) if a != b]
        assert len(mine) == len(ref) and not diff, (name, diff[:5])


def test_nhswps_32_4_field_dumps(run_32_4):
    n2 = 32 * 32
    for sub, names, n in (("3d", ("ql", "d", "g", "r", "w", "pn"), n2 * 5), ("2d", ("d", "zeta", "q", "g", "h"), n2)):
        for name in names:
            t, mine = read_r4(os.path.join(run_32_4, sub, name + ".r4"), n)
            tr, ref = read_r4(os.path.join(CASE, "complete", "32_4", sub, name + ".r4"), n)
            assert np.array_equal(t, tr), (sub, name)
            err = np.abs(mine - ref).max()
            print(f"{sub}/{name}.r4: max |GPU - reference| = {err:.1e}", flush=True)
            np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-8, err_msg=f"{sub}/{name}")


def test_nhswps_32_4_spectra_against_oracle(oracle, run_32_4):
    # the oracle's loop to the last save (t = 1.0), then its savegrid spectra, as spectra.asc prints them:
    # with tsim just below 1.0 the loop stops at t = 1.0 and saves after it (main_invert and psolve, as the
    # save inside the loop at t = 1.0 does before its last step)
    ng, nz = 32, 4
    raw = np.fromfile(os.path.join(CASE, "balinit", "32_4_sw_init.r8"), dtype="<f8").reshape(3, ng * ng + 1)
    s = oracle.Spectral(ng, nz)
    filt = s.get("filt")
    fields = {}
    for name, rec in zip(("qs", "ds", "gs"), raw[:, 1:]):
        spec = s.field()
        s.ptospc3d(np.tile(rec, nz + 1), spec, 0, nz)
        for iz in range(nz + 1):
            spec[iz * ng * ng] = 0.0
            spec[iz * ng * ng:(iz + 1) * ng * ng] *= filt
        fields[name] = spec
    o = oracle.State(s, **fields)
    rc, _, eco, _ = oracle.run(o, 1.0 - 0.5 / ng, 0.25)
    assert rc == 0 and len(eco) == 5 and eco[-1][0] == 1.0
    _, _, spec = oracle.savegrid(o)
    lines = open(os.path.join(run_32_4, "spectra.asc")).read().strip().split("\n")
    kmaxred = ng // 2
    assert len(lines) == 5 * (kmaxred + 1)
    last = lines[-(kmaxred + 1):]
    assert last[0].split() == ["1.000000", str(kmaxred)]
    got = np.array([[float(x) for x in ln.split()] for ln in last[1:]])
    k = np.arange(1, kmaxred + 1)
    np.testing.assert_allclose(got[:, 0], np.log10(k), rtol=0, atol=6e-9)
    # the shell sums themselves: to the 8 printed decimals of their log10 (relative 5e-8), except in shells
    # that hold nothing but rounding noise (below 1e-10 of the field's largest shell), where any two
    # summation orders differ
    for c, add in ((1, 0.0), (2, 1e-32), (3, 1e-32)):
        lin, ref = 10.0 ** got[:, c], spec[c - 1, k] + add
        bad = np.abs(lin - ref) > 5e-8 * ref + 1e-10 * ref.max()
        assert not bad.any(), (c, k[bad], lin[bad], ref[bad])


def test_savegrid_against_oracle_128x8(sw, oracle):
    ng, nz = 128, 8
    qs, ds = sw.synth_kminus3(ng, nz, seed=4, k0=8.0, delta_rms=0.01 * 4 * np.pi)
    o = oracle.State(oracle.Spectral(ng, nz), qs=qs, ds=ds, gs=np.zeros_like(qs), ggen=True)
    o.spectral.main_invert(o.qs, o.ds, o.gs, o.r, o.u, o.v, o.zeta)
    assert oracle.advance(o)[0] == 0
    assert oracle.psolve(o)[0] == 0
    g = sw.State(sw.Spectral(ng, nz), **{f: getattr(o, f) for f in ALL})
    a3, a2, aspec = sw.savegrid(g)
    b3, b2, bspec = oracle.savegrid(o)
    # the f32 fields agree to within one f32 rounding of the FP64 values they come from
    for a, b, what in ((a3, b3, "3d"), (a2, b2, "2d")):
        for i in range(a.shape[0]):
            tol = 1.2e-7 * np.abs(b[i]).max()
            err = np.abs(a[i].astype(np.float64) - b[i]).max()
            assert err <= tol, (what, i, err, tol)
    e = rel_l2(aspec, bspec)
    print(f"savegrid 128x8: spectra rel L2 {e:.1e}", flush=True)
    assert e < 1e-12


def test_checkpoint_round_trip(sw, oracle, tmp_path):
    ng, nz, t0 = 24, 4, 0.624_999_999_999_999_9
    p = f"nhswps/testdata/advance/{ng}_{nz}/"
    f = {n: gold(p + n + ".bin") for n in ALL}
    a = sw.State(sw.Spectral(ng, nz), t=t0, ggen=True, **f)
    path = str(tmp_path / "state.r8")
    a.save(path)
    assert os.path.getsize(path) == 8 + 16 * ng * ng * (nz + 1) * 8
    b = sw.State(sw.Spectral(ng, nz))
    b.load(path)
    b.set_time(b.t, True)
    assert b.t == a.t
    for n in ALL:
        assert np.array_equal(a.get(n), b.get(n)), n
    sw.advance(a)
    sw.advance(b)
    for n in ALL:
        assert np.array_equal(a.get(n), b.get(n)), n
