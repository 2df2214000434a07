"""shallow-water_amd: the nhswps per-timestep hot path on MI355X (gfx950).

A thin ctypes layer over libsw_hip.so (C ABI in include/shallow_water.h) that mirrors the reference
crate's API for this path: stafft (factorisen, initfft, forfft, revfft), sta2dfft (D2FFT with
ptospc, spctop, xderiv, yderiv), spectral (Spectral with main_invert, jacob, divs, ptospc3d,
spctop3d, deal3d, deal2d) and nhswps (State with vertical, coeffs, cpsource, psolve, source,
diagnose, advance) and balinit.  Arrays are float64 numpy arrays in the reference layout (element (i,j,iz) at
i + ng*j + ng*ng*iz).  Errors raise SwError carrying the library's sw_status code where the
reference would exit(1) or panic.

The library must be built (make -C shallow-water_amd, or __graft_entry__.build()); importing this
module without it raises ImportError: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsw_hip.so")

FP64, MIXED = 0, 1
PROF_OFF, PROF_STRAIGHT, PROF_PSOLVE_DERIV, PROF_PSOLVE_COMB, PROF_PSOLVE_TRI, PROF_STAGES = 0, 1, 2, 3, 4, 5
STAGES = ["main_invert", "psolve_setup", "psolve_iteration", "psolve_end", "source", "update", "diagnose"]
FIELDS = ["u", "v", "w", "z", "zx", "zy", "r", "ri", "aa", "zeta", "pn", "dpn", "ps", "qs", "ds", "gs"]
STATUS = {0: "SW_OK", -1: "SW_ERR_ARG", -2: "SW_ERR_PSOLVE_MAXITER", -3: "SW_ERR_PSOLVE_DIVERGED",
          -4: "SW_ERR_FACTOR", -5: "SW_ERR_UNSUPPORTED", -6: "SW_ERR_HIP", -7: "SW_ERR_NONFINITE"}


class SwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (run make -C shallow-water_amd)")
_lib = C.CDLL(LIB_PATH)
_lib.sw_last_error.restype = C.c_char_p
_lib.sw_spectral_get.restype = C.c_long
_dp = C.POINTER(C.c_double)
_sp = C.POINTER(C.c_size_t)
_vp = C.c_void_p


def _check(rc):
    if rc != 0:
        raise SwError(rc, _lib.sw_last_error().decode())
    return rc


def _d(a):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_dp)


def _sz(f):
    a = np.ascontiguousarray(f, dtype=np.uint64)
    return a, a.ctypes.data_as(_sp)


def lib_path():
    return LIB_PATH


def init(device=0):
    _check(_lib.sw_init(C.c_int(device)))


def sync():
    _check(_lib.sw_sync())


# ---------------------------------------------------------------- stafft (src/stafft/mod.rs)
def factorisen(n):
    f = np.zeros(5, np.uint64)
    _check(_lib.sw_factorisen(C.c_size_t(n), f.ctypes.data_as(_sp)))
    return [int(x) for x in f]


def initfft(n):
    f = np.zeros(5, np.uint64)
    trig = np.zeros(2 * n)
    _check(_lib.sw_initfft(C.c_size_t(n), f.ctypes.data_as(_sp), _d(trig)))
    return [int(x) for x in f], trig


def forfft(m, n, x, trig=None, factors=None):
    """In place: m real transforms of length n of x(m,n) (element (i,j) at i + m*j).  trig must be
    initfft(n)'s table (checked; SwError otherwise); None: that table, formed here."""
    if trig is None:
        trig = initfft(n)[1]
    fa, fp = _sz(factors) if factors is not None else (None, None)
    _check(_lib.sw_forfft(C.c_size_t(m), C.c_size_t(n), _d(x), _d(trig), fp))


def revfft(m, n, x, trig=None, factors=None):
    """In place: the inverse of forfft (same arguments)."""
    if trig is None:
        trig = initfft(n)[1]
    fa, fp = _sz(factors) if factors is not None else (None, None)
    _check(_lib.sw_revfft(C.c_size_t(m), C.c_size_t(n), _d(x), _d(trig), fp))


# ---------------------------------------------------------------- sta2dfft (src/sta2dfft.rs)
def fft_set_tile(n, bytes_, B, nt):
    """Tile (rows per block B, threads nt) of the block-per-rows passes of length n over `bytes_`-byte
    values, for engines created afterwards; B = 0 drops the override (sw_fft_set_tile)."""
    _check(_lib.sw_fft_set_tile(C.c_size_t(n), C.c_int(bytes_), C.c_int(B), C.c_int(nt)))


def fft_get_tile(n, bytes_):
    """(B, nt) that an engine created now uses for length n."""
    b, t = C.c_int(), C.c_int()
    _check(_lib.sw_fft_get_tile(C.c_size_t(n), C.c_int(bytes_), C.byref(b), C.byref(t)))
    return b.value, t.value


def fft_time_tile(n, bytes_, B, nt, nslab, reps):
    """Microseconds per forward + inverse 2-D transform of nslab slabs with tile (B, nt)."""
    us = C.c_double()
    _check(_lib.sw_fft_time_tile(C.c_size_t(n), C.c_int(bytes_), C.c_int(B), C.c_int(nt), C.c_int(nslab),
                                 C.c_int(reps), C.byref(us)))
    return us.value


class D2FFT:
    def __init__(self, nx, ny, lx=2 * np.pi, ly=2 * np.pi):
        self.nx, self.ny = nx, ny
        self.kx, self.ky = np.zeros(nx), np.zeros(ny)
        h = _vp()
        _check(_lib.sw_d2fft_create(C.c_size_t(nx), C.c_size_t(ny), C.c_double(lx), C.c_double(ly), _d(self.kx),
                                    _d(self.ky), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.sw_d2fft_destroy(self._h)
            self._h = None

    def ptospc(self, rvar, svar):
        _check(_lib.sw_ptospc(self._h, _d(rvar), _d(svar)))

    def spctop(self, svar, rvar):
        _check(_lib.sw_spctop(self._h, _d(svar), _d(rvar)))

    def xderiv(self, rkx, var, der):
        _check(_lib.sw_xderiv(self._h, _d(rkx), _d(var), _d(der)))

    def yderiv(self, rky, var, der):
        _check(_lib.sw_yderiv(self._h, _d(rky), _d(var), _d(der)))


# ---------------------------------------------------------------- spectral (src/spectral/mod.rs)
class Spectral:
    def __init__(self, ng, nz, precision=FP64):
        self.ng, self.nz, self.precision = ng, nz, precision
        h = _vp()
        _check(_lib.sw_spectral_create(C.c_size_t(ng), C.c_size_t(nz), C.c_int(precision), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.sw_spectral_destroy(self._h)
            self._h = None

    def field(self):
        return np.zeros(self.ng * self.ng * (self.nz + 1))

    def slab(self):
        return np.zeros(self.ng * self.ng)

    def get(self, name):
        n = _lib.sw_spectral_get(self._h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        _lib.sw_spectral_get(self._h, name.encode(), _d(out))
        return out

    def profile(self, mode=PROF_STRAIGHT):
        """Time one kind of FFT pass launch of this handle with HIP events (PROF_STRAIGHT: the straight
        pass of every 2-D transform; PROF_PSOLVE_DERIV: the transposing pass of psolve's theta-derivative
        transform; PROF_PSOLVE_COMB: psolve's fused inverse-transform / source / forward-transform launch;
        PROF_PSOLVE_TRI: psolve's fused forward pass and column solve); PROF_OFF stops."""
        _check(_lib.sw_profile(self._h, C.c_int(int(mode))))

    def profile_read(self):
        """(total ms, total algorithmic bytes, launches) of the timed FFT pass launches."""
        ms, by, n = C.c_double(), C.c_double(), C.c_long()
        _check(_lib.sw_profile_read(self._h, C.byref(ms), C.byref(by), C.byref(n)))
        return ms.value, by.value, n.value

    def profile_stages(self):
        """After profile(PROF_STAGES) and some advance steps: (ms per stage as a dict over STAGES, steps,
        psolve iterations)."""
        ms = np.zeros(7)
        steps, its = C.c_long(), C.c_long()
        _check(_lib.sw_profile_stages(self._h, _d(ms), C.byref(steps), C.byref(its)))
        return dict(zip(STAGES, ms.tolist())), steps.value, its.value

    def set_tolerance(self, toler):
        _check(_lib.sw_set_tolerance(self._h, C.c_double(toler)))

    def main_invert(self, qs, ds, gs, r, u, v, zeta):
        _check(_lib.sw_main_invert(self._h, _d(qs), _d(ds), _d(gs), _d(r), _d(u), _d(v), _d(zeta)))

    def jacob(self, aa, bb, cs):
        _check(_lib.sw_jacob(self._h, _d(aa), _d(bb), _d(cs)))

    def divs(self, aa, bb, cs):
        _check(_lib.sw_divs(self._h, _d(aa), _d(bb), _d(cs)))

    def ptospc3d(self, fp, fs, izbeg, izend):
        _check(_lib.sw_ptospc3d(self._h, _d(fp), _d(fs), C.c_size_t(izbeg), C.c_size_t(izend)))

    def spctop3d(self, fs, fp, izbeg, izend):
        _check(_lib.sw_spctop3d(self._h, _d(fs), _d(fp), C.c_size_t(izbeg), C.c_size_t(izend)))

    def deal3d(self, fp):
        _check(_lib.sw_deal3d(self._h, _d(fp)))

    def deal2d(self, fp):
        _check(_lib.sw_deal2d(self._h, _d(fp)))


# ---------------------------------------------------------------- nhswps (src/nhswps)
class State:
    """Device-resident counterpart of the reference State (src/nhswps/mod.rs:87-133)."""

    def __init__(self, spectral, t=0.0, ggen=False, **fields):
        self.spectral = spectral
        h = _vp()
        _check(_lib.sw_state_create(spectral._h, C.byref(h)))
        self._h = h
        for name, a in fields.items():
            self.set(name, a)
        self.set_time(t, ggen)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.sw_state_destroy(self._h)
            self._h = None

    def set(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        _check(_lib.sw_state_set(self._h, C.c_int(FIELDS.index(name)), _d(a)))

    def get(self, name):
        out = self.spectral.field()
        _check(_lib.sw_state_get(self._h, C.c_int(FIELDS.index(name)), _d(out)))
        return out

    def set_time(self, t, ggen=False):
        _check(_lib.sw_state_set_time(self._h, C.c_double(t), C.c_int(1 if ggen else 0)))

    @property
    def t(self):
        t = C.c_double()
        _check(_lib.sw_state_get_time(self._h, C.byref(t)))
        return t.value

    def save(self, path):
        """FP64 checkpoint: an 8-byte header holding t, then the 16 fields in State order."""
        _check(_lib.sw_state_save(self._h, os.fsencode(path)))

    def load(self, path):
        _check(_lib.sw_state_load(self._h, os.fsencode(path)))

    def __getattr__(self, name):
        if name in FIELDS:
            return self.get(name)
        raise AttributeError(name)


def vertical(st):
    _check(_lib.sw_vertical(st._h))


def coeffs(st, sigx, sigy, cpt1, cpt2):
    _check(_lib.sw_coeffs(st._h, _d(sigx), _d(sigy), _d(cpt1), _d(cpt2)))


def cpsource(st, sp0):
    _check(_lib.sw_cpsource(st._h, _d(sp0)))


def psolve(st, nitmax=100):
    """Returns (iterations, final errp)."""
    it, errp = C.c_int(), C.c_double()
    _check(_lib.sw_psolve(st._h, C.c_int(nitmax), C.byref(it), C.byref(errp)))
    return it.value, errp.value


def source(st, sqs, sds, sgs):
    _check(_lib.sw_source(st._h, _d(sqs), _d(sds), _d(sgs)))


def diagnose(st):
    out = np.zeros(5)
    _check(_lib.sw_diagnose(st._h, _d(out)))
    return out


def advance(st):
    """One timestep; returns (monitor values [t, zrms^2/2, zrms, zmax, umax], psolve iterations)."""
    diag = np.zeros(5)
    iters = np.zeros(3, np.int32)
    _check(_lib.sw_advance(st._h, _d(diag), iters.ctypes.data_as(C.POINTER(C.c_int))))
    return diag, [int(i) for i in iters]


def energy(st):
    """The ecomp.asc row of savegrid (src/nhswps/mod.rs:365-420): [t, 0, ekin, ekin, epot, etot]."""
    out = np.zeros(6)
    _check(_lib.sw_energy(st._h, _d(out)))
    return out


def run(st, tsim, tgsave, max_rows=None):
    """The nhswps time loop (src/nhswps/mod.rs:285-345) from the state's qs, ds, gs and time: returns
    (monitor rows (n, 5), ecomp rows (m, 6), total psolve iterations)."""
    dt = 1.0 / st.spectral.ng
    if max_rows is None:
        max_rows = int((tsim - st.t) / dt) + 4
    mon = np.zeros((max_rows, 5))
    eco = np.zeros((max_rows, 6))
    nm, ne, it = C.c_int(), C.c_int(), C.c_int()
    _check(_lib.sw_run(st._h, C.c_double(tsim), C.c_double(tgsave), _d(mon), C.c_int(max_rows), C.byref(nm), _d(eco),
                       C.c_int(max_rows), C.byref(ne), C.byref(it)))
    return mon[:min(nm.value, max_rows)], eco[:min(ne.value, max_rows)], it.value


def savegrid(st):
    """savegrid's output (src/nhswps/mod.rs:347-652): (3-D f32 fields (6, N): ql d g r w pn; 2-D f32 fields
    (5, ng^2): d zeta q g h; spectra (3, ng + 1): zeta, delta, gamma)."""
    ng, nz = st.spectral.ng, st.spectral.nz
    n2 = ng * ng
    d3 = np.zeros((6, n2 * (nz + 1)), np.float32)
    d2 = np.zeros((5, n2), np.float32)
    spec = np.zeros((3, ng + 1))
    _check(_lib.sw_savegrid(st._h, d3.ctypes.data_as(C.POINTER(C.c_float)), d2.ctypes.data_as(C.POINTER(C.c_float)),
                            _d(spec)))
    return d3, d2, spec


def read_r8(path, n):
    """n little-endian f64 values after an 8-byte header (the *_init.r8 layout)."""
    out = np.zeros(n)
    _check(_lib.sw_read_r8(os.fsencode(path), _d(out), C.c_size_t(n)))
    return out


def nhswps(out_dir, ng, nz, time_step, duration, save_interval, precision=FP64):
    """nhswps(&Parameters) (src/nhswps/mod.rs:150-345) on the GPU: reads out_dir/{qq,dd,gg}_init.r8, runs
    the time loop and writes monitor.asc, ecomp.asc, spectra.asc, 2d/*.r4 and 3d/*.r4 into out_dir.
    Returns the total psolve iterations."""
    it = C.c_int()
    _check(_lib.sw_nhswps(os.fsencode(out_dir), C.c_size_t(ng), C.c_size_t(nz), C.c_double(time_step),
                          C.c_double(duration), C.c_double(save_interval), C.c_int(precision), C.byref(it)))
    return it.value


def balinit(spectral, qq_init):
    """balinit (src/balinit.rs:21-436) on the GPU for one ng*ng PV anomaly: the balanced (qq, dd, gg)
    that sw_init.r8 holds, and the iteration count."""
    q0 = np.ascontiguousarray(qq_init, np.float64).ravel()
    if q0.size != spectral.ng * spectral.ng:
        raise ValueError("qq_init must hold ng*ng values")
    qq, dd, gg = spectral.slab(), spectral.slab(), spectral.slab()
    it = C.c_int()
    _check(_lib.sw_balinit(spectral._h, _d(q0), _d(qq), _d(dd), _d(gg), C.byref(it)))
    return qq, dd, gg, it.value


def balinit_file(out_dir, ng, nz):
    """balinit(&Parameters) as the program runs it: reads out_dir/qq_init.r8 and writes out_dir/sw_init.r8.
    Returns the iteration count."""
    it = C.c_int()
    _check(_lib.sw_balinit_file(os.fsencode(out_dir), C.c_size_t(ng), C.c_size_t(nz), C.byref(it)))
    return it.value


class Ensemble:
    """Member states advanced together (sw_ensemble_*): each member needs its own Spectral handle; members
    of one shape and precision run as batches; the batches and single members are spread over `nstreams`
    HIP streams (0: one per batch or single member, at most 8), each driven by one host thread."""

    def __init__(self, states, nstreams=0):
        self.states = list(states)
        arr = (C.c_void_p * len(self.states))(*[st._h.value for st in self.states])
        h = _vp()
        _check(_lib.sw_ensemble_create(C.c_int(len(self.states)), arr, C.c_int(nstreams), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.sw_ensemble_destroy(self._h)
            self._h = None

    def set_graphs(self, enable=True, margin=4):
        """Graph replay of the members' steps (on by default): see sw_ensemble_set_graphs."""
        _check(_lib.sw_ensemble_set_graphs(self._h, C.c_int(1 if enable else 0), C.c_int(margin)))

    def set_batching(self, max_members=32):
        """Same-shape members advanced as batches of at most max_members (0 or 1: off); see
        sw_ensemble_set_batching."""
        _check(_lib.sw_ensemble_set_batching(self._h, C.c_int(max_members)))

    def batches(self):
        """(number of batches, members in batches, batched steps so far, streams)."""
        a, b, c, d = C.c_int(), C.c_int(), C.c_long(), C.c_int()
        _check(_lib.sw_ensemble_batches(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return a.value, b.value, c.value, d.value

    def stats(self):
        """(graph replays, eager redos, captures) so far."""
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        _check(_lib.sw_ensemble_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def advance(self):
        """One timestep of every member; returns (monitor values (n, 5), psolve iterations (n, 3),
        status (n,)); raises SwError when a member failed (its status is the error code)."""
        n = len(self.states)
        diag = np.zeros((n, 5))
        iters = np.zeros((n, 3), np.int32)
        status = np.zeros(n, np.int32)
        _check(_lib.sw_ensemble_advance(self._h, _d(diag), iters.ctypes.data_as(C.POINTER(C.c_int)),
                                        status.ctypes.data_as(C.POINTER(C.c_int))))
        return diag, iters, status


def synth_kminus3(ng, nz, seed, k0=8.0, zeta_rms=0.5 * 4 * np.pi, delta_rms=0.0):
    """Seeded spectral PV (and divergence) with an isotropic k^-3 spectrum peaked at k0, identical on
    every level; returns (qs, ds) as spectral fields."""
    qs = np.zeros(ng * ng * (nz + 1))
    ds = np.zeros_like(qs)
    _check(_lib.sw_synth_kminus3(C.c_size_t(ng), C.c_size_t(nz), C.c_ulonglong(seed), C.c_double(k0),
                                 C.c_double(zeta_rms), C.c_double(delta_rms), _d(qs), _d(ds)))
    return qs, ds
