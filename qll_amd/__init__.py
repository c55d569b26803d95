"""qll_amd: MI355X-native QLL Dslash-and-solver path.

Python access to libqll_amd.so through its C ABI (include/qll_amd.h and the
QLL drop-in symbols of include/qll.h).  All arithmetic runs in the library's
HIP kernels; this module only marshals arguments.  Importing works without a
GPU; every call that needs one raises QllAmdError when no device or kernel
image is usable -- there is no host fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqll_amd.so")

F64, F32, F16, Q16 = 0, 1, 2, 3
EVEN, ODD = 0, 1
STAG, WILSON, STAG_FB = 0, 1, 2   # STAG_FB: forward links only (createStaggeredLinksFb)

# fat7 + Naik coefficients {one_link, three_staple, five_staple, seven_staple,
# lepage, naik} of the benchmark configurations
FAT7_ASQTAD = (5.0 / 8, -1.0 / 16, 1.0 / 64, -1.0 / 384, 0.0, -1.0 / 24)


class QllAmdError(RuntimeError):
    pass


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("reliable_updates", C.c_int),
                ("true_rsq", C.c_double), ("iter_rsq", C.c_double),
                ("seconds", C.c_double), ("converged", C.c_int),
                ("restarts", C.c_int), ("escalations", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class QllLayout(C.Structure):
    """qamd_qll_layout: the fields of a QLL Layout the transfers need."""
    _fields_ = [("physGeom", C.c_int * 4), ("innerGeom", C.c_int * 4),
                ("outerGeom", C.c_int * 4), ("nSitesOuter", C.c_int),
                ("nSitesInner", C.c_int), ("nEvenOuter", C.c_int),
                ("innerCb", C.c_int), ("innerCbDir", C.c_int)]


_lib = None
_vp = C.c_void_p
_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    """Load libqll_amd.so (built by qll_amd.build); raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QllAmdError("libqll_amd.so is not built (run qll_amd.build.build())")
        L = C.CDLL(LIB_PATH)
        sigs = {
            "qamd_last_error": (C.c_char_p, []),
            "qamd_init": (C.c_int, []),
            "qamd_sync": (C.c_int, []),
            "qamd_device_name": (C.c_int, [C.c_char_p, C.c_int]),
            "qamd_lattice_create": (_vp, [C.POINTER(C.c_int)]),
            "qamd_lattice_destroy": (None, [_vp]),
            "qamd_field_create": (_vp, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "qamd_field_destroy": (None, [_vp]),
            "qamd_field_zero": (C.c_int, [_vp]),
            "qamd_field_copy": (C.c_int, [_vp, _vp]),
            "qamd_field_from_lex": (C.c_int, [_vp, _d]),
            "qamd_field_to_lex": (C.c_int, [_vp, _d]),
            "qamd_field_from_qll": (C.c_int, [_vp, C.POINTER(QllLayout), _vp, C.c_int]),
            "qamd_field_to_qll": (C.c_int, [_vp, C.POINTER(QllLayout), _vp, C.c_int]),
            "qamd_field_gaussian": (C.c_int, [_vp, C.c_uint64, C.c_uint64]),
            "qamd_field_norm2": (C.c_int, [_vp, C.POINTER(C.c_double)]),
            "qamd_field_redot": (C.c_int, [_vp, _vp, C.POINTER(C.c_double)]),
            "qamd_field_axpby": (C.c_int, [C.c_double, _vp, C.c_double, _vp]),
            "qamd_gauge_create": (_vp, [_vp]),
            "qamd_gauge_destroy": (None, [_vp]),
            "qamd_gauge_from_lex": (C.c_int, [_vp, _d]),
            "qamd_gauge_to_lex": (C.c_int, [_vp, _d]),
            "qamd_gauge_random_su3": (C.c_int, [_vp, C.c_uint64]),
            "qamd_gauge_stag_phases": (C.c_int, [_vp]),
            "qamd_fat7": (C.c_int, [_vp, _vp, _d, _vp, _vp]),
            "qamd_gauge_plaquette": (C.c_int, [_vp, _d]),
            "qamd_links_create": (_vp, [_vp, C.c_int, C.c_int, C.c_int]),
            "qamd_links_destroy": (None, [_vp]),
            "qamd_links_from_lex": (C.c_int, [_vp, _d]),
            "qamd_links_from_qll": (C.c_int, [_vp, C.POINTER(QllLayout), C.POINTER(_vp), C.POINTER(_vp), C.c_int]),
            "qamd_links_from_gauge": (C.c_int, [_vp, _vp, _vp]),
            "qamd_links_convert": (C.c_int, [_vp, _vp]),
            "qamd_links_reconstructs": (C.c_int, [_vp]),
            "qamd_links_bytes": (C.c_size_t, [_vp]),
            "qamd_dslash": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp, C.c_double, C.POINTER(C.c_double)]),
            "qamd_stag_normal": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.POINTER(C.c_double)]),
            "qamd_stag_cg": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.POINTER(SolveInfo)]),
            "qamd_stag_cg_mixed": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_double,
                                             C.POINTER(SolveInfo)]),
            "qamd_set_tuning": (C.c_int, [C.c_int]),
            "qamd_set_mall_share": (C.c_int, [C.c_int, C.c_int]),
            "qamd_release_cache": (C.c_int, []),
            "qamd_stag_sloppy_error": (C.c_int, [_vp, _vp, C.c_double, C.POINTER(C.c_double)]),
            "qamd_stag_cg_mixed_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                                   C.POINTER(SolveInfo)]),
            "qamd_stag_cg_block": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.POINTER(SolveInfo)]),
            "qamd_stag_cg_mixed_block": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_double,
                                             C.POINTER(SolveInfo)]),
            "qamd_stag_cgms": (C.c_int, [_vp, C.POINTER(_vp), _vp, C.c_double, _d, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(SolveInfo)]),
            "qamd_stag_cgms_mixed": (C.c_int, [_vp, _vp, C.POINTER(_vp), _vp, C.c_double, _d, C.c_int, C.c_double,
                                               C.c_int, C.POINTER(SolveInfo)]),
            "qamd_time_dslash": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_double)]),
            "qamd_stream_copy": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "qamd_wilson_cgne": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_int,
                                           C.POINTER(SolveInfo)]),
            "qamd_field_gamma5": (C.c_int, [_vp]),
            "qamd_field_from_sites": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int64, _vp, C.c_int]),
            "qamd_field_to_sites": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int64, _vp, C.c_int]),
            "qamd_gauge_from_sites": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int64, C.POINTER(_vp), C.c_int,
                                                C.c_int]),
            "qamd_gauge_to_sites": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int64, C.POINTER(_vp), C.c_int, C.c_int]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise QllAmdError(lib().qamd_last_error().decode())


def _ptr(p):
    if not p:
        raise QllAmdError(lib().qamd_last_error().decode())
    return p


def init():
    _check(lib().qamd_init())


def sync():
    _check(lib().qamd_sync())


def device_name():
    buf = C.create_string_buffer(256)
    _check(lib().qamd_device_name(buf, 256))
    return buf.value.decode()


class Lattice:
    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.h = _ptr(lib().qamd_lattice_create((C.c_int * 4)(*self.dims)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qamd_lattice_destroy(self.h)
            self.h = None


class Field:
    """One parity of a colour-vector (ncomp=3) or spinor (ncomp=12) field."""

    def __init__(self, lat, prec=F64, ncomp=3, nrhs=1, parity=EVEN):
        self.lat, self.prec, self.ncomp, self.nrhs, self.parity = lat, prec, ncomp, nrhs, parity
        self.h = _ptr(lib().qamd_field_create(lat.h, prec, ncomp, nrhs, parity))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qamd_field_destroy(self.h)
            self.h = None

    def set_lex(self, a):
        """a: (nrhs, V, 2*ncomp) or (V, 2*ncomp) lexicographic; own parity read."""
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.nrhs * self.lat.V * 2 * self.ncomp
        _check(lib().qamd_field_from_lex(self.h, a))
        return self

    def get_lex(self, base=None):
        shape = (self.nrhs, self.lat.V, 2 * self.ncomp)
        out = np.zeros(shape) if base is None else np.ascontiguousarray(base, dtype=np.float64).reshape(shape).copy()
        _check(lib().qamd_field_to_lex(self.h, out.reshape(-1)))
        return out if self.nrhs > 1 else out[0]

    def gaussian(self, seed, stream):
        _check(lib().qamd_field_gaussian(self.h, seed, stream))
        return self

    def zero(self):
        _check(lib().qamd_field_zero(self.h))
        return self

    def copy_from(self, other):
        _check(lib().qamd_field_copy(self.h, other.h))
        return self

    def norm2(self):
        out = (C.c_double * self.nrhs)()
        _check(lib().qamd_field_norm2(self.h, out))
        return list(out) if self.nrhs > 1 else out[0]

    def redot(self, other):
        out = (C.c_double * self.nrhs)()
        _check(lib().qamd_field_redot(self.h, other.h, out))
        return list(out) if self.nrhs > 1 else out[0]

    def axpby(self, a, x, b):
        """self = a*x + b*self"""
        _check(lib().qamd_field_axpby(a, x.h, b, self.h))
        return self

    def set_sites(self, coords, a):
        """a: (nsites, 2*ncomp) in the site order of coords (nsites, 4)."""
        c = np.ascontiguousarray(coords, dtype=np.int32)
        a = np.ascontiguousarray(a)
        prec = F64 if a.dtype == np.float64 else F32
        assert a.dtype in (np.float64, np.float32) and a.size == c.shape[0] * 2 * self.ncomp
        _check(lib().qamd_field_from_sites(self.h, c.ctypes.data_as(C.POINTER(C.c_int)), c.shape[0],
                                           a.ctypes.data, prec))
        return self

    def get_sites(self, coords, base=None, dtype=np.float64):
        c = np.ascontiguousarray(coords, dtype=np.int32)
        shape = (c.shape[0], 2 * self.ncomp)
        out = np.zeros(shape, dtype=dtype) if base is None else np.array(base, dtype=dtype).reshape(shape)
        _check(lib().qamd_field_to_sites(self.h, c.ctypes.data_as(C.POINTER(C.c_int)), c.shape[0], out.ctypes.data,
                                         F64 if dtype == np.float64 else F32))
        return out


class Gauge:
    def __init__(self, lat):
        self.lat = lat
        self.h = _ptr(lib().qamd_gauge_create(lat.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qamd_gauge_destroy(self.h)
            self.h = None

    def set_lex(self, u):
        _check(lib().qamd_gauge_from_lex(self.h, np.ascontiguousarray(u, dtype=np.float64).reshape(-1)))
        return self

    def get_lex(self):
        u = np.empty((4, self.lat.V, 18))
        _check(lib().qamd_gauge_to_lex(self.h, u.reshape(-1)))
        return u

    def random_su3(self, seed):
        _check(lib().qamd_gauge_random_su3(self.h, seed))
        return self

    def set_sites(self, coords, u, row_major=False):
        """u: (4, nsites, 18) in the site order of coords (nsites, 4)."""
        c = np.ascontiguousarray(coords, dtype=np.int32)
        arrs = [np.ascontiguousarray(u[mu]) for mu in range(4)]
        prec = F64 if arrs[0].dtype == np.float64 else F32
        ptrs = (_vp * 4)(*[a.ctypes.data for a in arrs])
        _check(lib().qamd_gauge_from_sites(self.h, c.ctypes.data_as(C.POINTER(C.c_int)), c.shape[0], ptrs, prec,
                                           1 if row_major else 0))
        return self

    def get_sites(self, coords, row_major=False):
        c = np.ascontiguousarray(coords, dtype=np.int32)
        out = np.zeros((4, c.shape[0], 18))
        ptrs = (_vp * 4)(*[out[mu].ctypes.data for mu in range(4)])
        _check(lib().qamd_gauge_to_sites(self.h, c.ctypes.data_as(C.POINTER(C.c_int)), c.shape[0], ptrs, F64,
                                         1 if row_major else 0))
        return out

    def stag_phases(self):
        _check(lib().qamd_gauge_stag_phases(self.h))
        return self

    def plaquette(self):
        p = np.empty(6)
        _check(lib().qamd_gauge_plaquette(self.h, p))
        return p


def fat7(u, coef=FAT7_ASQTAD, ulong=None, want_long=True):
    fl = Gauge(u.lat)
    ll = Gauge(u.lat) if want_long else None
    _check(lib().qamd_fat7(fl.h, ll.h if ll else None, np.asarray(coef, dtype=np.float64), u.h,
                           (ulong or u).h))
    return fl, ll


class Links:
    def __init__(self, lat, kind=STAG, prec=F64, nterms=16):
        self.lat, self.kind, self.prec, self.nterms = lat, kind, prec, nterms
        self.h = _ptr(lib().qamd_links_create(lat.h, kind, prec, nterms))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qamd_links_destroy(self.h)
            self.h = None

    def set_lex(self, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        assert links.shape == (self.nterms, self.lat.V, 18)
        _check(lib().qamd_links_from_lex(self.h, links.reshape(-1)))
        return self

    def from_gauge(self, fat, lng=None):
        _check(lib().qamd_links_from_gauge(self.h, fat.h, lng.h if lng is not None else None))
        return self

    @property
    def nbytes(self):
        """Device bytes of the link set."""
        return int(lib().qamd_links_bytes(self.h))

    @property
    def reconstructs(self):
        """FP64 Dslash rebuilds the long terms' third rows (fewer bytes)."""
        return bool(lib().qamd_links_reconstructs(self.h))

    def converted(self, prec):
        out = Links(self.lat, self.kind, prec, self.nterms)
        _check(lib().qamd_links_convert(out.h, self.h))
        return out


def dslash(links, out, inp, a=0.0, w=None, b=1.0, want_norm=False):
    n = C.c_double()
    _check(lib().qamd_dslash(links.h, out.h, inp.h, a, w.h if w is not None else None, b,
                             C.byref(n) if want_norm else None))
    return n.value if want_norm else None


def stag_normal(links, Ap, p, tmp, mass):
    v = C.c_double()
    _check(lib().qamd_stag_normal(links.h, Ap.h, p.h, tmp.h, mass, C.byref(v)))
    return v.value


def stag_cg(links, x, b, mass, rsq, maxits):
    info = SolveInfo()
    _check(lib().qamd_stag_cg(links.h, x.h, b.h, mass, rsq, maxits, C.byref(info)))
    return info.as_dict()


def stag_cg_mixed(l64, ls, x, b, mass, rsq, maxits, delta=0.1):
    info = SolveInfo()
    _check(lib().qamd_stag_cg_mixed(l64.h, ls.h, x.h, b.h, mass, rsq, maxits, delta, C.byref(info)))
    return info.as_dict()


def stag_cg_block(links, x, b, mass, rsq, maxits):
    """Plain FP64 block CG on every rhs of the even block b; one info dict per rhs."""
    infos = (SolveInfo * b.nrhs)()
    _check(lib().qamd_stag_cg_block(links.h, x.h, b.h, mass, rsq, maxits, infos))
    return [i.as_dict() for i in infos]


def stag_cg_mixed_block(l64, ls, x, b, mass, rsq, maxits, delta=0.1):
    """Solve every rhs of the even block b together; one info dict per rhs."""
    infos = (SolveInfo * b.nrhs)()
    _check(lib().qamd_stag_cg_mixed_block(l64.h, ls.h, x.h, b.h, mass, rsq, maxits, delta, infos))
    return [i.as_dict() for i in infos]


def set_tuning(mode):
    """Launch-shape autotuning: True/1 on (default), False/0 off, k >= 2 forces
    candidate k - 2 (tests)."""
    _check(lib().qamd_set_tuning(int(mode)))


def release_cache():
    """Free the device blocks the library keeps for reuse."""
    _check(lib().qamd_release_cache())


def set_mall_share(mb64=-1, mb=-1):
    """Megabytes of each staggered link set (FP64; FP32 and 16-bit) kept in
    the MALL from call to call; 0 streams every link load, -1 the default."""
    _check(lib().qamd_set_mall_share(int(mb64), int(mb)))


def stag_sloppy_error(l64, ls, mass):
    """|(A_sloppy - A) v| / |v| of the mixed solvers' sloppy normal operator."""
    e = C.c_double()
    _check(lib().qamd_stag_sloppy_error(l64.h, ls.h, mass, C.byref(e)))
    return e.value


class _Problem(C.Structure):
    _fields_ = [("l64", _vp), ("ls", _vp), ("x", _vp), ("b", _vp), ("mass", C.c_double)]


def stag_cg_mixed_batch(problems, rsq, maxits, delta=0.1, use_graph=True):
    """problems: list of (l64, ls, x, b, mass).  Solves them together; returns
    one info dict per rhs, problems in order."""
    arr = (_Problem * len(problems))(*[_Problem(p[0].h, p[1].h, p[2].h, p[3].h, p[4]) for p in problems])
    n = sum(p[3].nrhs for p in problems)
    infos = (SolveInfo * n)()
    _check(lib().qamd_stag_cg_mixed_batch(C.cast(arr, C.c_void_p), len(problems), rsq, maxits, delta,
                                          1 if use_graph else 0, infos))
    return [i.as_dict() for i in infos]


def stag_cgms(links, xs, b, mass, shifts, rsq, maxits):
    info = SolveInfo()
    arr = (_vp * len(xs))(*[x.h for x in xs])
    _check(lib().qamd_stag_cgms(links.h, arr, b.h, mass, np.asarray(shifts, dtype=np.float64), len(xs), rsq,
                                maxits, C.byref(info)))
    return info.as_dict()


def stag_cgms_mixed(l64, ls, xs, b, mass, shifts, rsq, maxits):
    infos = (SolveInfo * len(xs))()
    arr = (_vp * len(xs))(*[x.h for x in xs])
    _check(lib().qamd_stag_cgms_mixed(l64.h, ls.h, arr, b.h, mass, np.asarray(shifts, dtype=np.float64), len(xs),
                                      rsq, maxits, infos))
    return [i.as_dict() for i in infos]


def wilson_cgne(l64, x, b, kappa, rsq, maxits, ls=None):
    """Solve (1 - kappa^2 H_eo H_oe) x = b for every rhs of the block b."""
    infos = (SolveInfo * b.nrhs)()
    _check(lib().qamd_wilson_cgne(l64.h, ls.h if ls is not None else None, x.h, b.h, kappa, rsq, maxits,
                                  infos))
    return [i.as_dict() for i in infos]


def stream_copy(nbytes, nrep=20):
    """Measured copy bandwidth (read + write) in GB/s and ms per copy."""
    g, ms = C.c_double(), C.c_double()
    _check(lib().qamd_stream_copy(nbytes, nrep, C.byref(g), C.byref(ms)))
    return g.value, ms.value


def time_dslash(links, out, inp, nrep):
    ms = C.c_double()
    _check(lib().qamd_time_dslash(links.h, out.h, inp.h, nrep, C.byref(ms)))
    return ms.value


def asqtad_links_device(lat, seed, coef=FAT7_ASQTAD, prec=F64, kind=STAG):
    """Seeded SU(3) -> staggered phases -> fat7/Naik -> 16-term link set (or
    the 8-term forward-only set, kind=STAG_FB), all on the GPU."""
    u = Gauge(lat).random_su3(seed).stag_phases()
    fl, ll = fat7(u, coef)
    l64 = Links(lat, kind, F64, 16 if kind == STAG else 8).from_gauge(fl, ll)
    del fl, ll, u
    return l64 if prec == F64 else l64.converted(prec)
