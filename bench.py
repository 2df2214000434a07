#!/usr/bin/env python3
"""Throughput of the nhswps timestep (src/nhswps/advance.rs) on one MI355X.

A step is one full `advance` (predictor + two correctors: 3x main_invert, 3x psolve, 3x source)
with all fields resident in HBM, on a synthetic seeded k^-3 PV state at ng=512, nz=16.  The FP64 step
and the mixed-precision step (psolve's coefficient fields and fixed-point iteration in FP32, its
tridiagonal and residual in FP64, every other stage and the state in FP64) are both timed, and the
mixed step's per-field relative L2 error against the FP64 step after one step is measured (q, delta,
gamma, h, p).  The headline `value` is the faster mode that keeps every field within 1e-5 of the FP64
step: the mixed step if it does, else the FP64 step itself.  A point-step is one stored level point
advanced one step: ng^2 (nz+1) per step.  The mixed mode's iteration stops at 1e-6 (the FP32
transforms' floor), the FP64 one at the reference's 1e-9, so the FP64 step is also timed at 1e-6
(equal tolerance).

Also reported: per-stage times per step and the time per psolve iteration for both modes (HIP events
around the stages); the roofline of the largest of psolve's three kernels (HIP events on the library
stream around its launches; algorithmic bytes per launch as documented for each); and a CPU baseline:
the oracle (C++ restatement of the reference algorithm, OpenMP over levels, built -march=native on
this host when a compiler is there) for one warm-up and three timed steps of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--ng 512] [--nz 16]
       python bench.py --large 200   (the ng=1536, nz=24 workload alone: mixed and FP64, 200 steps)
       python bench.py --ensemble-only   (the 64-member ensemble workload alone)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "point-steps/s @ng512 nz16: mixed vs FP64 GPU at rel-L2≤1e-5; % HBM roofline"
HBM_PEAK_GBS = 8000.0
ACCURACY_BAR = 1e-5  # relative L2 against the FP64 step after one step, per field (q, delta, gamma, h, p)
FIELDS = ["u", "v", "w", "z", "zx", "zy", "r", "ri", "aa", "zeta", "pn", "dpn", "ps", "qs", "ds", "gs"]


def load_sw():
    import importlib.util
    pkg = os.path.join(ROOT, "shallow-water_amd")
    spec = importlib.util.spec_from_file_location("shallow_water_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["shallow_water_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def initial_fields(sw, ng, nz, seed):
    """Seeded k^-3 PV (k0 = 8, rms zeta = f/2) on identical levels, delta = gamma = 0, pn = 0, then
    the loader's inversion (src/nhswps/mod.rs:238-277) done on the GPU in FP64."""
    return seeded_init(sw, ng, nz, seed, 0.5)


def new_state(sw, ng, nz, precision, init):
    spec = sw.Spectral(ng, nz, precision)
    st = sw.State(spec, t=0.0, ggen=True, **init)
    return spec, st


def time_steps(sw, spec, st, steps, warmup, prof_modes):
    for _ in range(warmup):
        st.set_time(st.t, True)
        sw.advance(st)
    sw.sync()
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st.set_time(st.t, True)
        _, it = sw.advance(st)
        iters.append(sum(it))
    sw.sync()
    dt = time.perf_counter() - t0
    # per-stage times over two more steps
    spec.profile(sw.PROF_STAGES)
    for _ in range(2):
        st.set_time(st.t, True)
        sw.advance(st)
    stage_ms, nsteps, nits = spec.profile_stages()
    spec.profile(sw.PROF_OFF)
    stages = {"ms_per_step": {k: v / nsteps for k, v in stage_ms.items()},
              "psolve_iterations_per_step": nits / nsteps,
              "ms_per_psolve_iteration": stage_ms["psolve_iteration"] / max(nits, 1)}
    # the same workload again, per profiled kernel, with HIP events on the library stream around each
    # launch of that kernel
    prof = {}
    for name, mode in prof_modes.items():
        spec.profile(mode)
        nstep = min(steps, 2)
        for _ in range(nstep):
            st.set_time(st.t, True)
            sw.advance(st)
        ms, nbytes, launches = spec.profile_read()
        spec.profile(sw.PROF_OFF)
        prof[name] = dict(ms_per_step=ms / nstep, bytes=nbytes, launches=launches, ms=ms)
    return dt / steps * 1e3, iters, prof, stages


def measured_traffic(site, kernel, ng, nz, precision):
    """HBM bytes per launch of a profiled kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    readings (profiles/hbm_traffic_<site>_ng<ng>_nz<nz>_<precision>.txt), scaled by the copy-kernel
    calibration; None when no reading for this kernel and workload is on file."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{site}_ng{ng}_nz{nz}_{precision}.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        k, _, v = line.strip().partition(" ")
        kv[k] = v
    if kv.get("kernel") != kernel or kv.get("workload") != f"ng={ng} nz={nz} precision={precision}":
        return None
    kib = 1024.0
    return (float(kv["fetch_kib_per_dispatch"]) * float(kv["fetch_scale"])
            + float(kv["write_kib_per_dispatch"]) * float(kv["write_scale"])) * kib


def kernel_names(ng, nz, precision):
    """The profiled psolve kernels on the fused path (transform lengths 128/256/512, nz 4/8/16/24/32): the
    derivative rows of the first guess (once per psolve call), and per iteration the source fold (tiled
    derivative inputs) and the column solve that also forms the next iteration's derivative rows."""
    t = "double" if precision == "fp64" else "float"
    lv, nd = nz + 1, 1 + 3 * nz
    nw = 4 if precision == "fp64" else 8  # waves per fold block; tile rows 2 nw
    return {"psolve_deriv": f"fftw_rows_cols_inv<{t}, {ng}, 8, {lv}, {nd}, {2 * nw}, load-only: first-guess derivative rows>",
            "psolve_comb": f"fftw_inv_comb_fwd<{t}, {ng}, {nw}, psolve source fold, tiled inputs>",
            "psolve_tri": f"fftw_rows_cols_inv<{t}, {ng}, 8, {lv}, {nd}, {2 * nw}, psolve column solve + derivative rows>"}


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def one_step_errors(sw, ng, nz, init):
    out = {}
    for p in (sw.FP64, sw.MIXED):
        spec, st = new_state(sw, ng, nz, p, init)
        sw.advance(st)
        out[p] = {f: st.get(f) for f in ("qs", "ds", "gs", "z", "pn")}
    n2 = ng * ng
    err = {}
    for name, f in (("q", "qs"), ("delta", "ds"), ("gamma", "gs"), ("p", "pn")):
        err[name] = rel_l2(out[sw.MIXED][f], out[sw.FP64][f])
    err["h"] = rel_l2(out[sw.MIXED]["z"][nz * n2:], out[sw.FP64]["z"][nz * n2:])
    return err


def seeded_init(sw, ng, nz, seed, zeta_frac, delta_frac=0.0):
    """initial_fields with rms zeta = zeta_frac * f (and an independent random delta of rms delta_frac * f)."""
    qs, ds = sw.synth_kminus3(ng, nz, seed=seed, k0=8.0, zeta_rms=zeta_frac * 4 * np.pi,
                              delta_rms=delta_frac * 4 * np.pi)
    gs = np.zeros_like(qs)
    spec = sw.Spectral(ng, nz, sw.FP64)
    r, u, v, zeta = spec.field(), spec.field(), spec.field(), spec.field()
    spec.main_invert(qs, ds, gs, r, u, v, zeta)
    return dict(qs=qs, ds=ds, gs=gs, r=r, u=u, v=v, zeta=zeta)


def run_steps(sw, ng, nz, precision, init, steps):
    """Advance a fresh state up to `steps` steps (the reference's time loop with ggen set on every step,
    src/nhswps/mod.rs:289-324).  Returns (completed steps, fields of the last completed step or None,
    the error that stopped the run or None, FP64 monitor values)."""
    spec, st = new_state(sw, ng, nz, precision, init)
    done, stop, diag = 0, None, None
    for _ in range(steps):
        st.set_time(st.t, True)
        try:
            d, _ = sw.advance(st)
        except sw.SwError as e:
            stop = str(e)
            break
        diag = [float(x) for x in d]
        done += 1
    fields = {f: st.get(f) for f in ("qs", "ds", "gs", "z", "pn")} if stop is None else None
    del st, spec
    return done, fields, stop, diag


def long_run_errors(sw, ng, nz, init, steps):
    """Per-field relative L2 error of the mixed-precision state against the FP64 state after `steps`
    steps from the same initial state, or after the last step both complete when the pressure
    iteration stops converging first (psolve's 100-iteration cap: the reference exits there,
    src/nhswps/psolve.rs:360-366)."""
    target = steps
    done, f64, stop, diag = run_steps(sw, ng, nz, sw.FP64, init, target)
    stopped = None
    if stop is not None:
        stopped = f"fp64 step {done + 1}: {stop}"
        target = done
        done, f64, _, diag = run_steps(sw, ng, nz, sw.FP64, init, target)
    mdone, mix, mstop, _ = run_steps(sw, ng, nz, sw.MIXED, init, target)
    if mstop is not None:
        stopped = (stopped + "; " if stopped else "") + f"mixed step {mdone + 1}: {mstop}"
        target = mdone
        _, f64, _, diag = run_steps(sw, ng, nz, sw.FP64, init, target)
        _, mix, _, _ = run_steps(sw, ng, nz, sw.MIXED, init, target)
    n2 = ng * ng
    err = {}
    for name, f in (("q", "qs"), ("delta", "ds"), ("gamma", "gs"), ("p", "pn")):
        err[name] = rel_l2(mix[f], f64[f])
    err["h"] = rel_l2(mix["z"][nz * n2:], f64["z"][nz * n2:])
    return {"steps_requested": steps, "steps_compared": target, "stopped": stopped, "rel_l2_mixed_vs_fp64": err,
            "fp64_monitor_t_zrms2half_zrms_zmax_umax": diag}


def large_bench(sw, steps, seed=1, ng=1536, nz=24, zeta_frac=0.5):
    """The large mixed-radix workload: ng = 1536 (2^9 3), nz = 24, seeded k^-3 PV (k0 = 8, rms zeta =
    zeta_frac f) plus an independent random delta of rms 0.01 f, `steps` steps in the mixed mode and in FP64
    from the same state (each step timed; a run stops early where psolve's iteration stops converging, as
    the reference exits there, src/nhswps/psolve.rs:360-366).  Reports ms per step and point-steps/s of each
    mode over the steps it completed, psolve iterations, and the mixed-vs-FP64 errors after 1 step and after
    the last step when both modes completed every step."""
    init = seeded_init(sw, ng, nz, seed, zeta_frac, 0.01)
    points = ng * ng * (nz + 1)
    out = {"workload": f"nhswps advance, ng={ng}, nz={nz}, seeded k^-3 PV (k0=8, rms zeta={zeta_frac} f) + random "
                       f"delta (rms 0.01 f)", "steps_requested": steps}
    fields = {}
    for p, name in ((sw.MIXED, "mixed"), (sw.FP64, "fp64")):
        spec, st = new_state(sw, ng, nz, p, init)
        its, times, stop, snap1 = [], [], None, None
        for k in range(steps):
            st.set_time(st.t, True)
            t0 = time.perf_counter()
            try:
                _, it = sw.advance(st)
                sw.sync()
            except sw.SwError as e:
                stop = f"step {k + 1}: {e}"
                break
            times.append(time.perf_counter() - t0)
            its.append(sum(it))
            if k == 0:
                snap1 = {f: st.get(f) for f in ("qs", "ds", "gs", "z", "pn")}
        fields[name] = (snap1, {f: st.get(f) for f in ("qs", "ds", "gs", "z", "pn")} if stop is None else None,
                        len(times))
        timed = times[1:] if len(times) > 1 else times  # the first step holds one-off set-up
        ms = 1e3 * sum(timed) / max(len(timed), 1)
        out[name] = {"steps_completed": len(times), "stopped": stop, "ms_per_step": ms,
                     "point_steps_per_s": points / (ms * 1e-3) if ms > 0 else None,
                     "psolve_iters_per_step_first_last": [its[:3], its[-3:]]}
        print(f"[bench] large {name}: {out[name]}", file=sys.stderr, flush=True)
        del st, spec
    n2 = ng * ng

    def errs(a, b):
        e = {k: rel_l2(a[f], b[f]) for k, f in (("q", "qs"), ("delta", "ds"), ("gamma", "gs"), ("p", "pn"))}
        e["h"] = rel_l2(a["z"][nz * n2:], b["z"][nz * n2:])
        return e

    out["rel_l2_mixed_vs_fp64_after_1_step"] = errs(fields["mixed"][0], fields["fp64"][0])
    if fields["mixed"][1] is not None and fields["fp64"][1] is not None:
        out[f"rel_l2_mixed_vs_fp64_after_{steps}_steps"] = errs(fields["mixed"][1], fields["fp64"][1])
    out["mixed_over_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    return out


def ensemble_bench(sw, steps, warmup, nmember=64, runs=((0, "batched"), (1, "batched"), (4, "eager"))):
    """The batched-ensemble workload: nmember members, member m with ng = (64, 128, 256)[m % 3],
    nz = (4, 8)[(m // 3) % 2], k0 = (4, 8, 16)[(m // 6) % 3], seed 1000 + m, seeded k^-3 PV with rms
    zeta = 0.3 f (at f/2 the reference algorithm's pressure iteration diverges for the k0 = 4, ng = 64
    members), mixed precision (FP32 psolve coefficient fields and iteration, FP64 elsewhere), each
    member with its own spectral handle, advanced together by sw_ensemble_advance on 1 or 4 streams:
    "batched" = the members of each shape (about 11) advanced as one batch (every launch of the step covers
    all of them), "eager" = one member at a time per stream, "graphs" = one member at a time, replayed
    from a recorded hipGraph.  Returns the aggregate point-steps/s per run."""
    shapes = [((64, 128, 256)[m % 3], (4, 8)[(m // 3) % 2], (4.0, 8.0, 16.0)[(m // 6) % 3], 1000 + m)
              for m in range(nmember)]
    points = sum(ng * ng * (nz + 1) for ng, nz, _, _ in shapes)
    out = {"members": nmember, "points_per_step": points, "precision": "mixed", "rms_zeta": "0.3 f",
           "shapes": "ng=(64,128,256)[m%3], nz=(4,8)[(m//3)%2], k0=(4,8,16)[(m//6)%3], seed=1000+m"}
    for ns, mode in runs:
        states = []
        for ng, nz, k0, seed in shapes:
            qs, ds = sw.synth_kminus3(ng, nz, seed=seed, k0=k0, zeta_rms=0.3 * 4 * np.pi)
            gs = np.zeros_like(qs)
            s64 = sw.Spectral(ng, nz, sw.FP64)
            r, u, v, zeta = s64.field(), s64.field(), s64.field(), s64.field()
            s64.main_invert(qs, ds, gs, r, u, v, zeta)
            del s64
            states.append(sw.State(sw.Spectral(ng, nz, sw.MIXED), ggen=True, qs=qs, ds=ds, gs=gs, r=r, u=u, v=v,
                                   zeta=zeta))
        ens = sw.Ensemble(states, nstreams=ns)
        ens.set_batching(32 if mode == "batched" else 0)
        ens.set_graphs(mode == "graphs")

        def step():
            for st in states:
                st.set_time(st.t, True)
            ens.advance()

        for _ in range(warmup):
            step()
        sw.sync()
        r0 = ens.stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        dt = (time.perf_counter() - t0) / steps
        r1 = ens.stats()
        nst = ens.batches()[3]
        key = f"streams_{nst}_{mode}"
        out[key] = {"ms_per_step": dt * 1e3, "point_steps_per_s": points / dt}
        if mode == "graphs":
            out[key].update(replays=r1[0] - r0[0], redos=r1[1] - r0[1], captures=r1[2] - r0[2])
        if mode == "batched":
            out[key]["batches"] = ens.batches()[0]
        print(f"[bench] ensemble of {nmember} on {nst} stream(s), {mode}: {dt * 1e3:.2f} ms/step, "
              f"{points / dt:.4g} point-steps/s, {out[key]}", file=sys.stderr, flush=True)
        del ens, states
    return out


def hbm_ceiling():
    """What plain torch read (sum), write (fill) and copy kernels reach on 1 GiB FP64 buffers, beyond the
    256 MB last-level cache (HIP events, 20 launches each): the practical ceilings of HBM traffic with
    those read/write mixes, beside the 8 TB/s peak that roofline.frac is quoted against."""
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    out = {}
    for name, fn, moved in (("read", lambda: a.sum(), n * 8), ("write", lambda: b.fill_(2.0), n * 8),
                            ("copy", lambda: b.copy_(a), 2 * n * 8)):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            fn()
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 20 * 1e-3
        out[name] = {"TB_per_s": moved / t / 1e12, "frac_of_peak": moved / t / (HBM_PEAK_GBS * 1e9)}
    del a, b
    torch.cuda.empty_cache()
    return out


def transform_rate(sw, ng=512, nslab=256, reps=20):
    """The stand-alone FP64 2-D transforms (ptospc3d, spctop3d) of nslab slabs of ng^2 held in device memory
    (537 MB in and out at the defaults, beyond the last-level cache): TB/s over both passes, each pass
    reading and writing every value."""
    import ctypes as C
    import torch
    n = ng * ng * nslab
    spec = sw.Spectral(ng, nslab - 1)
    a = torch.randn(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    out = {}
    for name in ("ptospc3d", "spctop3d"):
        fn = getattr(sw._lib, "sw_" + name)
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t]
        for _ in range(2):
            assert fn(spec._h, a.data_ptr(), b.data_ptr(), 0, nslab - 1) == 0
        sw.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(spec._h, a.data_ptr(), b.data_ptr(), 0, nslab - 1)
        sw.sync()
        t = (time.perf_counter() - t0) / reps
        out[name] = {"us": t * 1e6, "TB_per_s": 4 * n * 8 / t / 1e12}
    del spec, a, b
    torch.cuda.empty_cache()
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sw, ng, nz, init, timed=3):
    """The oracle (reference algorithm, C++ restatement) on this host: one warm-up step, then `timed`
    steps, each from the previous (ggen set, as the GPU run).  Built -march=native on this host when
    possible (make -C oracle native), else the prebuilt x86-64-v3 library.  Threads: the host's CPU share
    for one GPU (16) or its core count if smaller."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import subprocess
    import _oracle
    build = "-O3 -march=x86-64-v3 (prebuilt)"
    try:
        subprocess.run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"), "native"], check=True, timeout=300,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        _oracle.LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_sw_native.so")
        build = "-O3 -march=native (built on this host)"
    except Exception as e:  # no compiler or a failed build: time the prebuilt library
        print(f"[bench] oracle native build unavailable ({e}); timing the prebuilt library", file=sys.stderr)
    nproc = os.cpu_count() or 1
    cores = _oracle.set_threads(min(nproc, 16))
    o = _oracle.State(_oracle.Spectral(ng, nz), ggen=True, **init)
    rc, _, _ = _oracle.advance(o)
    if rc != 0:
        return None
    t0 = time.perf_counter()
    for _ in range(timed):
        o.ggen = True
        rc, _, _ = _oracle.advance(o)
        if rc != 0:
            return None
    dt = (time.perf_counter() - t0) / timed
    return {"value": ng * ng * (nz + 1) / dt, "unit": "point-steps/s", "cores": cores, "kind": "port",
            "nproc": nproc, "cpu_model": cpu_model(), "build": build, "ms_per_step": dt * 1e3,
            "sample": f"{timed} advance steps after 1 warm-up step at ng={ng}, nz={nz} (FP64, {dt:.1f} s per step), "
                      f"same seeded state; reference algorithm, C++ restatement, {cores} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ng", type=int, default=512)
    ap.add_argument("--nz", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=3,
                    help="timed steps of the 64-member ensemble workload reported beside the headline (0: skip)")
    ap.add_argument("--long-steps", type=int, default=1000,
                    help="steps of the mixed-vs-FP64 accuracy run reported beside the one-step errors (0: skip)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the 64-member ensemble workload (--ensemble-steps steps on --ensemble-streams streams)")
    ap.add_argument("--ensemble-streams", type=int, default=0, help="0: one per batch or single member")
    ap.add_argument("--ensemble-mode", choices=("batched", "eager", "graphs"), default="batched")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="only advance the state --steps steps in this precision after --warmup steps, with no "
                         "report (the run that rocprofv3 profiles)")
    ap.add_argument("--large", type=int, default=0, metavar="STEPS",
                    help="only the large mixed-radix workload (ng=1536, nz=24, k^-3 PV + random delta): STEPS steps "
                         "mixed and FP64, errors at step 1 and at the end")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="only measure what plain read, write and copy kernels reach on 1 GiB buffers (beyond the "
                         "256 MB last-level cache): the practical HBM ceilings beside the 8 TB/s peak")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (twice) and exit: the reference reading "
                         "for the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 2 copies of {args.copy_calibration} MiB (read {args.copy_calibration} MiB "
              f"+ write {args.copy_calibration} MiB each)", file=sys.stderr, flush=True)
        return
    if args.hbm_ceiling:
        print(json.dumps({"hbm_ceiling_1GiB": hbm_ceiling(), "transform_pair_537MB": transform_rate(load_sw())}),
              flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (ensemble members, or the "
              "levels of one member with an all-to-all per psolve iteration, would shard across GPUs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # the HBM ceilings first: torch initialises its own HIP context before the library does
    ceiling = hbm_ceiling() if not (args.large or args.ensemble_only or args.only) else None
    sw = load_sw()
    if args.large:
        # the bench state (rms zeta = f/2), on which the reference algorithm's pressure iteration stops
        # converging within the run, then rms zeta = 0.3 f, on which it completes every step
        print(json.dumps({"large": {"rms_zeta_0.5f": large_bench(sw, args.large, args.seed),
                                    "rms_zeta_0.3f": large_bench(sw, args.large, args.seed, zeta_frac=0.3)}}),
              flush=True)
        return
    if args.ensemble_only:
        ens = ensemble_bench(sw, args.ensemble_steps, 2, runs=((args.ensemble_streams, args.ensemble_mode),))
        print(json.dumps({"ensemble": ens}), flush=True)
        return
    ng, nz = args.ng, args.nz
    points = ng * ng * (nz + 1)
    init = initial_fields(sw, ng, nz, args.seed)
    if args.only:
        spec, st = new_state(sw, ng, nz, sw.FP64 if args.only == "fp64" else sw.MIXED, init)
        for k in range(args.warmup + args.steps):
            st.set_time(st.t, True)
            _, it = sw.advance(st)
            print(f"[bench] {args.only} step {k}: psolve iterations {it}", file=sys.stderr, flush=True)
        sw.sync()
        return

    res = {}
    modes = {"psolve_deriv": sw.PROF_PSOLVE_DERIV, "psolve_comb": sw.PROF_PSOLVE_COMB, "psolve_tri": sw.PROF_PSOLVE_TRI}
    for p, name, tol in ((sw.FP64, "fp64", None), (sw.MIXED, "mixed", None), (sw.FP64, "fp64_tol_1e-6", 1e-6)):
        spec, st = new_state(sw, ng, nz, p, init)
        if tol is not None:
            spec.set_tolerance(tol)
        ms, iters, prof, stages = time_steps(sw, spec, st, args.steps, args.warmup, modes if tol is None else {})
        res[name] = dict(ms=ms, iters=iters, prof=prof, stages=stages, value=points / (ms * 1e-3))
        print(f"[bench] {name}: {ms:.3f} ms/step, {points / (ms * 1e-3):.4g} point-steps/s, psolve iterations "
              f"per step {iters}; stages {stages}", file=sys.stderr, flush=True)
        del st, spec
    err = one_step_errors(sw, ng, nz, init)
    print(f"[bench] mixed vs fp64 after 1 step: {err}", file=sys.stderr, flush=True)
    long_runs = None
    if args.long_steps > 0:
        # the bench state (rms zeta = f/2), then a milder one (rms zeta = 0.3 f) on which the reference
        # algorithm's pressure iteration converges for the whole run
        long_runs = {"rms_zeta_0.5f": long_run_errors(sw, ng, nz, init, args.long_steps),
                     "rms_zeta_0.3f": long_run_errors(sw, ng, nz, seeded_init(sw, ng, nz, args.seed, 0.3),
                                                      args.long_steps)}
        for k, v in long_runs.items():
            print(f"[bench] long run {k}: {v}", file=sys.stderr, flush=True)

    ens = ensemble_bench(sw, args.ensemble_steps, 2) if args.ensemble_steps > 0 else None

    # headline: the fastest precision mode that meets the accuracy bar against the FP64 step
    meets = all(v <= ACCURACY_BAR for v in err.values())
    head = "mixed" if meets else "fp64"
    h, f64, m, f64t = res[head], res["fp64"], res["mixed"], res["fp64_tol_1e-6"]
    # roofline of the dominant profiled kernel (largest share of the step)
    site = max(h["prof"], key=lambda k: h["prof"][k]["ms_per_step"])
    pk = h["prof"][site]
    achieved = pk["bytes"] / (pk["ms"] * 1e-3) / 1e9 if pk["ms"] > 0 else 0.0
    names = kernel_names(ng, nz, head)
    kname = names[site]
    others = {k: {"kernel": names[k], "ms_per_step": v["ms_per_step"], "avg_launch_us": v["ms"] * 1e3 / max(v["launches"], 1),
                  "achieved_GBs": v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0}
              for k, v in h["prof"].items()}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sw, ng, nz, init)
    line = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"nhswps advance (3x main_invert, psolve, source), ng={ng}, nz={nz}, seeded k^-3 PV "
                               f"(k0=8, rms zeta=f/2), delta=gamma=0", "ng": ng, "nz": nz,
                   "precision": head},
        "headline_precision": head,
        "accuracy_bar": {"rel_l2_max": ACCURACY_BAR, "mixed_meets_bar": meets,
                         "rel_l2_mixed_vs_fp64_after_1_step": err,
                         "long_runs": long_runs},
        "fp64": {"value": f64["value"], "ms_per_step": f64["ms"], "psolve_iters_per_step": f64["iters"],
                 "stages": f64["stages"]},
        "mixed": {"value": m["value"], "ms_per_step": m["ms"], "psolve_iters_per_step": m["iters"],
                  "dtype": "f64 state and stages; f32 psolve coefficient fields and iteration (tolerance 1e-6), "
                           "f64 tridiagonal, residual and reductions",
                  "stages": m["stages"],
                  "over_fp64": m["value"] / f64["value"],
                  "over_fp64_equal_tolerance": m["value"] / f64t["value"]},
        "fp64_tol_1e-6": {"value": f64t["value"], "ms_per_step": f64t["ms"], "psolve_iters_per_step": f64t["iters"],
                          "stages": f64t["stages"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(site, kname, ng, nz, head),
                     "kernel": kname,
                     "launches": pk["launches"], "avg_launch_us": pk["ms"] * 1e3 / max(pk["launches"], 1),
                     "algorithmic_bytes_per_launch": pk["bytes"] / max(pk["launches"], 1),
                     "profiled_kernels": others,
                     "hbm_ceiling_1GiB": ceiling},
        "ensemble": ens,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
