"""GPU checks of the ensemble entry point (sw_ensemble_*): members of different ng, nz and precision
advanced together on several streams give exactly the states of the same members advanced one at a
time with sw_advance (same kernels, same data, deterministic reductions), and one ensemble step
equals the oracle's step for a member (the oracle being the parity reference of sw_advance)."""
import numpy as np
import pytest
from conftest import rel_l2
from _tolerances import STEP_TOL

pytestmark = pytest.mark.gpu
ALL = ["ri", "r", "u", "v", "w", "zeta", "z", "zx", "zy", "ps", "pn", "dpn", "aa", "qs", "ds", "gs"]
MEMBERS = [(64, 4, "fp64", 4.0, 11), (128, 4, "mixed", 8.0, 12), (96, 6, "fp64", 16.0, 13), (64, 8, "mixed", 16.0, 14),
           (256, 8, "fp64", 8.0, 15), (128, 8, "fp64", 4.0, 16)]


def init(sw, ng, nz, k0, seed):
    # rms zeta = 0.3 f: at f/2 the reference algorithm's pressure iteration diverges for some of these
    # members (64x4 with k0 = 4: the oracle stops with the same error), and the mixed mode loses 64x8
    # with k0 = 16
    qs, ds = sw.synth_kminus3(ng, nz, seed=seed, k0=k0, zeta_rms=0.3 * 4 * np.pi)
    gs = np.zeros_like(qs)
    s = sw.Spectral(ng, nz, sw.FP64)
    r, u, v, zeta = s.field(), s.field(), s.field(), s.field()
    s.main_invert(qs, ds, gs, r, u, v, zeta)
    return dict(qs=qs, ds=ds, gs=gs, r=r, u=u, v=v, zeta=zeta)


def states(sw):
    out = []
    for ng, nz, prec, k0, seed in MEMBERS:
        p = sw.FP64 if prec == "fp64" else sw.MIXED
        out.append(sw.State(sw.Spectral(ng, nz, p), ggen=True, **init(sw, ng, nz, k0, seed)))
    return out


def test_ensemble_equals_members_advanced_alone(sw):
    a, b = states(sw), states(sw)
    ens = sw.Ensemble(a, nstreams=3)
    for _ in range(2):
        diag, iters, status = ens.advance()
        assert not status.any()
        for i, st in enumerate(b):
            st.set_time(st.t, True)
            d, it = sw.advance(st)
            assert np.array_equal(diag[i], d) and list(iters[i]) == it, i
        for st in a:
            st.set_time(st.t, True)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.t == y.t
        for f in ALL:
            assert np.array_equal(x.get(f), y.get(f)), (i, f)


@pytest.mark.parametrize("margin", [0, 2])
def test_ensemble_graph_replays_equal_members_alone(sw, margin):
    # six steps: the first eager (and recorded), the rest graph replays; with margin 0 a psolve that needs
    # one more iteration than the last step overruns its captured budget, so the undo-and-redo path runs too
    a, b = states(sw), states(sw)
    ens = sw.Ensemble(a, nstreams=2)
    ens.set_graphs(True, margin)
    for step in range(6):
        diag, iters, status = ens.advance()
        assert not status.any()
        for i, st in enumerate(b):
            st.set_time(st.t, True)
            d, it = sw.advance(st)
            assert np.array_equal(diag[i], d) and list(iters[i]) == it, (step, i)
        for st in a:
            st.set_time(st.t, True)
    replays, redos, captures = ens.stats()
    print(f"margin {margin}: {replays} replays, {redos} redos, {captures} captures", flush=True)
    # members with nz = 6 have no device-controlled psolve and always run eagerly
    assert replays > 0
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.t == y.t
        for f in ALL:
            assert np.array_equal(x.get(f), y.get(f)), (i, f)


def test_ensemble_member_matches_oracle(sw, oracle):
    ng, nz, k0, seed = 96, 6, 8.0, 21
    f = init(sw, ng, nz, k0, seed)
    g = [sw.State(sw.Spectral(ng, nz), ggen=True, **f), sw.State(sw.Spectral(64, 4), ggen=True, **init(sw, 64, 4, 4.0, 22))]
    o = oracle.State(oracle.Spectral(ng, nz), ggen=True, **{k: g[0].get(k) for k in ALL})
    _, iters, status = sw.Ensemble(g, nstreams=2).advance()
    rc, _, iters2 = oracle.advance(o)
    assert rc == 0 and not status.any() and list(iters[0]) == iters2
    for fld in ALL:
        assert rel_l2(g[0].get(fld), getattr(o, fld)) < STEP_TOL[fld], fld


def test_ensemble_rejects_shared_spectral_handle(sw):
    s = sw.Spectral(64, 4)
    f = init(sw, 64, 4, 4.0, 1)
    with pytest.raises(sw.SwError) as e:
        sw.Ensemble([sw.State(s, ggen=True, **f), sw.State(s, ggen=True, **f)])
    assert e.value.code == -1


# same-shape members that the ensemble advances as batches (wave-kernel and block-per-rows transform
# lengths, FP64 and mixed), plus a member of its own shape
BATCHED = ([(128, 8, "fp64", k0, 30 + j) for j, k0 in enumerate((4.0, 8.0, 16.0))] +
           [(128, 4, "mixed", k0, 40 + j) for j, k0 in enumerate((8.0, 16.0, 8.0))] +
           [(96, 8, "fp64", k0, 50 + j) for j, k0 in enumerate((8.0, 16.0))] +
           [(64, 4, "mixed", 16.0, 60 + j) for j in range(2)] +
           [(256, 16, "fp64", 8.0, 70 + j) for j in range(2)] +
           [(64, 6, "fp64", 8.0, 80)])


def batched_states(sw):
    out = []
    for ng, nz, prec, k0, seed in BATCHED:
        p = sw.FP64 if prec == "fp64" else sw.MIXED
        out.append(sw.State(sw.Spectral(ng, nz, p), ggen=True, **init(sw, ng, nz, k0, seed)))
    return out


@pytest.mark.parametrize("batch_max", [32, 2])
def test_batched_ensemble_equals_members_alone(sw, batch_max):
    # three steps; at the third, member 0 runs with ggen off, so its batch runs its members one by one
    # batch_max 32: one stream per batch or single member (the default); 2: three streams
    a, b = batched_states(sw), batched_states(sw)
    ens = sw.Ensemble(a, nstreams=0 if batch_max == 32 else 3)
    ens.set_batching(batch_max)
    nb, nm, _, nst = ens.batches()
    print(f"batch_max {batch_max}: {nb} batches holding {nm} members, {nst} streams", flush=True)
    assert nm == len(BATCHED) - 1 and nb == (5 if batch_max == 32 else 7)
    assert nst == (nb + 1 if batch_max == 32 else 3)
    for step in range(3):
        ggen = [not (step == 2 and i == 0) for i in range(len(a))]
        for st, g in zip(a, ggen):
            st.set_time(st.t, g)
        diag, iters, status = ens.advance()
        assert not status.any()
        for i, st in enumerate(b):
            st.set_time(st.t, ggen[i])
            d, it = sw.advance(st)
            assert np.array_equal(diag[i], d) and list(iters[i]) == it, (step, i, diag[i], d, iters[i], it)
    assert ens.batches()[2] == 3 * nb - 1
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.t == y.t
        for f in ALL:
            assert np.array_equal(x.get(f), y.get(f)), (i, f)
