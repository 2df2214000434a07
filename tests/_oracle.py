"""ctypes binding of the CPU oracle (oracle/liboracle_sw.so), used only by the tests and by
bench.py's cpu_baseline leg.  The oracle restates the reference crate's algorithm; see
oracle/sw_oracle.h for the layout and return codes."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_sw.so")
FIELDS = ["u", "v", "w", "z", "zx", "zy", "r", "ri", "aa", "zeta", "pn", "dpn", "ps", "qs", "ds", "gs"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_spectral_new.restype = C.c_void_p
        _lib.orc_spectral_get.restype = C.c_long
        for name in ("orc_spectral_free", "orc_spectral_get", "orc_main_invert", "orc_jacob", "orc_divs",
                     "orc_ptospc3d", "orc_spctop3d", "orc_deal3d", "orc_deal2d", "orc_spec1d",
                     "orc_vertical", "orc_coeffs", "orc_cpsource", "orc_psolve", "orc_source",
                     "orc_diagnose", "orc_advance", "orc_energy", "orc_run", "orc_savegrid", "orc_balinit"):
            getattr(_lib, name).argtypes = None
    return _lib


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def szp(a):
    assert a.dtype == np.uint64
    return a.ctypes.data_as(C.POINTER(C.c_size_t))


def factorisen(n):
    f = np.zeros(5, np.uint64)
    rc = lib().orc_factorisen(C.c_size_t(n), szp(f))
    return rc, [int(x) for x in f]


def initfft(n):
    f = np.zeros(5, np.uint64)
    trig = np.zeros(2 * n)
    rc = lib().orc_initfft(C.c_size_t(n), szp(f), dp(trig))
    assert rc == 0
    return f, trig


def forfft(m, n, x, trig, factors):
    f = np.asarray(factors, np.uint64)
    return lib().orc_forfft(C.c_size_t(m), C.c_size_t(n), dp(x), dp(trig), szp(f))


def revfft(m, n, x, trig, factors):
    f = np.asarray(factors, np.uint64)
    return lib().orc_revfft(C.c_size_t(m), C.c_size_t(n), dp(x), dp(trig), szp(f))


def init2dfft(nx, ny, lx, ly):
    xf, yf = np.zeros(5, np.uint64), np.zeros(5, np.uint64)
    xt, yt, kx, ky = np.zeros(2 * nx), np.zeros(2 * ny), np.zeros(nx), np.zeros(ny)
    rc = lib().orc_init2dfft(C.c_size_t(nx), C.c_size_t(ny), C.c_double(lx), C.c_double(ly), szp(xf), szp(yf),
                             dp(xt), dp(yt), dp(kx), dp(ky))
    assert rc == 0
    return xf, yf, xt, yt, kx, ky


def ptospc(nx, ny, xf, yf, xt, yt, rvar, svar):
    lib().orc_ptospc(C.c_size_t(nx), C.c_size_t(ny), szp(np.asarray(xf, np.uint64)),
                     szp(np.asarray(yf, np.uint64)), dp(xt), dp(yt), dp(rvar), dp(svar))


def spctop(nx, ny, xf, yf, xt, yt, svar, rvar):
    lib().orc_spctop(C.c_size_t(nx), C.c_size_t(ny), szp(np.asarray(xf, np.uint64)),
                     szp(np.asarray(yf, np.uint64)), dp(xt), dp(yt), dp(svar), dp(rvar))


def xderiv(nx, ny, rkx, var, der):
    lib().orc_xderiv(C.c_size_t(nx), C.c_size_t(ny), dp(rkx), dp(var), dp(der))


def yderiv(nx, ny, rky, var, der):
    lib().orc_yderiv(C.c_size_t(nx), C.c_size_t(ny), dp(rky), dp(var), dp(der))


def set_threads(n):
    return lib().orc_set_threads(C.c_int(n))


class Spectral:
    """Oracle counterpart of the reference's Spectral (src/spectral/mod.rs:17-62)."""

    def __init__(self, ng, nz):
        self.ng, self.nz = ng, nz
        self.h = lib().orc_spectral_new(C.c_size_t(ng), C.c_size_t(nz))
        if not self.h:
            raise ValueError(f"oracle Spectral::new({ng},{nz}) failed")
        self._h = C.c_void_p(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_spectral_free(C.c_void_p(self.h))
            self.h = None

    def get(self, name):
        n = lib().orc_spectral_get(self._h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().orc_spectral_get(self._h, name.encode(), dp(out))
        return out

    def field(self):
        return np.zeros(self.ng * self.ng * (self.nz + 1))

    def slab(self):
        return np.zeros(self.ng * self.ng)

    def main_invert(self, qs, ds, gs, r, u, v, zeta):
        lib().orc_main_invert(self._h, dp(qs), dp(ds), dp(gs), dp(r), dp(u), dp(v), dp(zeta))

    def jacob(self, aa, bb, cs):
        lib().orc_jacob(self._h, dp(aa), dp(bb), dp(cs))

    def divs(self, aa, bb, cs):
        lib().orc_divs(self._h, dp(aa), dp(bb), dp(cs))

    def ptospc3d(self, fp, fs, izbeg, izend):
        lib().orc_ptospc3d(self._h, dp(fp), dp(fs), C.c_size_t(izbeg), C.c_size_t(izend))

    def spctop3d(self, fs, fp, izbeg, izend):
        lib().orc_spctop3d(self._h, dp(fs), dp(fp), C.c_size_t(izbeg), C.c_size_t(izend))

    def deal3d(self, fp):
        lib().orc_deal3d(self._h, dp(fp))

    def deal2d(self, fp):
        lib().orc_deal2d(self._h, dp(fp))

    def spec1d(self, ss, spec):
        lib().orc_spec1d(self._h, dp(ss), dp(spec))

    def balinit(self, qq_init):
        """src/balinit.rs:21-436 on one ng*ng PV anomaly: (qq, dd, gg, iterations)"""
        q0 = np.ascontiguousarray(qq_init, np.float64).ravel()
        assert q0.size == self.ng * self.ng
        qq, dd, gg = self.slab(), self.slab(), self.slab()
        it = lib().orc_balinit(self._h, dp(q0), dp(qq), dp(dd), dp(gg))
        return qq, dd, gg, it


class State:
    """Oracle counterpart of the reference's State (src/nhswps/mod.rs:87-133)."""

    def __init__(self, spectral, t=0.0, ggen=False, **fields):
        self.spectral = spectral
        n = spectral.ng * spectral.ng * (spectral.nz + 1)
        for name in FIELDS:
            a = fields.get(name)
            setattr(self, name, np.zeros(n) if a is None else np.array(a, np.float64).ravel(order="K").copy())
        self.t = t
        self.ggen = ggen

    def _ptrs(self):
        arr = (C.POINTER(C.c_double) * 16)(*[dp(getattr(self, f)) for f in FIELDS])
        return arr

    def copy(self):
        return State(self.spectral, self.t, self.ggen, **{f: getattr(self, f) for f in FIELDS})


def vertical(st):
    return lib().orc_vertical(st.spectral._h, st._ptrs())


def coeffs(st, sigx, sigy, cpt1, cpt2):
    return lib().orc_coeffs(st.spectral._h, st._ptrs(), dp(sigx), dp(sigy), dp(cpt1), dp(cpt2))


def cpsource(st, sp0):
    return lib().orc_cpsource(st.spectral._h, st._ptrs(), dp(sp0))


def psolve(st, nitmax=100):
    it, errp = C.c_int(0), C.c_double(0)
    rc = lib().orc_psolve(st.spectral._h, st._ptrs(), C.c_int(nitmax), C.byref(it), C.byref(errp))
    return rc, it.value, errp.value


def source(st, sqs, sds, sgs):
    return lib().orc_source(st.spectral._h, st._ptrs(), dp(sqs), dp(sds), dp(sgs))


def diagnose(st):
    out = np.zeros(5)
    lib().orc_diagnose(st.spectral._h, st._ptrs(), C.c_double(st.t), dp(out))
    return out


def advance(st, nitmax=100):
    t = C.c_double(st.t)
    diag = np.zeros(5)
    iters = np.zeros(3, np.int32)
    rc = lib().orc_advance(st.spectral._h, st._ptrs(), C.byref(t), C.c_int(1 if st.ggen else 0), C.c_int(nitmax),
                           dp(diag), iters.ctypes.data_as(C.POINTER(C.c_int)))
    st.t = t.value
    return rc, diag, [int(i) for i in iters]


def energy(st):
    out = np.zeros(6)
    lib().orc_energy(st.spectral._h, st._ptrs(), C.c_double(st.t), dp(out))
    return out


def savegrid(st):
    """savegrid's output as numbers: (3-D f32 fields (6, N): ql d g r w pn; 2-D f32 fields (5, ng^2):
    d zeta q g h; spectra (3, ng + 1): zeta, delta, gamma)."""
    ng, nz = st.spectral.ng, st.spectral.nz
    n2 = ng * ng
    d3 = np.zeros((6, n2 * (nz + 1)), np.float32)
    d2 = np.zeros((5, n2), np.float32)
    spec = np.zeros((3, ng + 1))
    lib().orc_savegrid(st.spectral._h, st._ptrs(), d3.ctypes.data_as(C.POINTER(C.c_float)),
                       d2.ctypes.data_as(C.POINTER(C.c_float)), dp(spec))
    return d3, d2, spec


def run(st, tsim, tgsave, nitmax=100, max_rows=None):
    dt = 1.0 / st.spectral.ng
    if max_rows is None:
        max_rows = int((tsim - st.t) / dt) + 4
    mon = np.zeros((max_rows, 5))
    eco = np.zeros((max_rows, 6))
    t = C.c_double(st.t)
    nm, ne, it = C.c_int(), C.c_int(), C.c_int()
    rc = lib().orc_run(st.spectral._h, st._ptrs(), C.byref(t), C.c_double(tsim), C.c_double(tgsave), C.c_int(nitmax),
                       dp(mon), C.c_int(max_rows), C.byref(nm), dp(eco), C.c_int(max_rows), C.byref(ne), C.byref(it))
    st.t = t.value
    return rc, mon[:min(nm.value, max_rows)], eco[:min(ne.value, max_rows)], it.value
