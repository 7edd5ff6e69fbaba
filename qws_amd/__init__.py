"""qws_amd -- MI355X (gfx950) Wilson-clover domain-decomposed solver.

The product is the C-ABI library qws_amd/lib/libqws_mi355x.so (HIP kernels and
the C++ host solvers, declared in include/qws_mi355x.h).  This module is a thin
ctypes binding of its device-resident interface (qwsg_*), used by the tests and
the benchmark.  There is no CPU fallback: when the library is not built, or a
call fails, an exception is raised.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("QWS_MI355X_LIB", os.path.join(HERE, "lib", "libqws_mi355x.so"))
HEADER = os.path.join(os.path.dirname(HERE), "include", "qws_mi355x.h")

F64, F32, F16 = 0, 1, 2
HOPS_ALL, HOPS_INSIDE, HOPS_FACE = 0, 1, 2

_lib = None


class QwsError(RuntimeError):
    pass


class SolveParams(ctypes.Structure):
    _fields_ = [("nsap", ctypes.c_int), ("nm", ctypes.c_int), ("tol", ctypes.c_double), ("maxiter", ctypes.c_int),
                ("tol_s", ctypes.c_double), ("maxiter_s", ctypes.c_int), ("true_tol", ctypes.c_double),
                ("max_restarts", ctypes.c_int), ("sap_precision", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("outer_iters", ctypes.c_int), ("inner_iters", ctypes.c_int), ("restarts", ctypes.c_int),
                ("true_residual", ctypes.c_double), ("seconds", ctypes.c_double), ("sap_escalated", ctypes.c_int)]


def lib():
    """Load libqws_mi355x.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise QwsError(f"{LIBPATH} is not built: run `python -m qws_amd.build`")
        L = ctypes.CDLL(LIBPATH)
        vp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
        L.qwsg_last_error.restype = ctypes.c_char_p
        L.qwsg_create.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.qwsg_destroy.argtypes = [vp]
        L.qwsg_set_kappa.argtypes = [vp, ctypes.c_double]
        L.qwsg_set_sap_fused.argtypes = [vp, ctypes.c_int]
        L.qwsg_set_sap_face_buffer.argtypes = [vp, ctypes.c_int]
        L.qwsg_set_sap_source_update.argtypes = [vp, ctypes.c_int]
        L.qwsg_volume.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
        for n in ("qwsg_load_gauge", "qwsg_load_clover", "qwsg_load_gauge_f32", "qwsg_load_clover_f32"):
            getattr(L, n).argtypes = [vp, vp]
        L.qwsg_spinor_alloc.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.qwsg_spinor_free.argtypes = [vp]
        L.qwsg_spinor_upload.argtypes = [vp, ctypes.c_int, vp, vp]
        L.qwsg_spinor_download.argtypes = [vp, ctypes.c_int, vp, vp]
        L.qwsg_spinor_zero.argtypes = [vp, ctypes.c_int, vp]
        L.qwsg_dslash.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int]
        L.qwsg_face.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_double, ctypes.c_int]
        L.qwsg_clover.argtypes = [vp, ctypes.c_int, vp, vp]
        L.qwsg_jinv.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int]
        L.qwsg_prec_ddd.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int]
        L.qwsg_prec_s.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int]
        L.qwsg_prec_ddd_noalloc.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int]
        L.qwsg_bicgstab_precdd.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ip]
        L.qwsg_msap_h.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int]
        L.qwsg_server_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(SolveParams), ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)]
        L.qwsg_server_run.argtypes = [vp, vp, ctypes.c_int, vp]
        L.qwsg_server_destroy.argtypes = [vp]
        L.qwsg_eo_load_gauge.argtypes = [vp, vp]
        L.qwsg_eo_load_clover.argtypes = [vp, vp]
        L.qwsg_eo_from_lattice.argtypes = [vp]
        L.qwsg_eo_spinor_alloc.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.qwsg_eo_spinor_upload.argtypes = [vp, ctypes.c_int, vp, vp]
        L.qwsg_eo_spinor_download.argtypes = [vp, ctypes.c_int, vp, vp]
        L.qwsg_deo.argtypes = [vp] + [ctypes.c_int] * 6 + [ctypes.c_double, vp, vp, vp]
        L.qwsg_mtilde.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp]
        L.qwsg_eo_bicgstab.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ip]
        L.qwsg_eo_cg.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ip]
        L.qwsg_eo_mcg.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, ip]
        L.qwsg_bicgstab_precdd_h.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
        L.qwsg_solve.argtypes = [vp, vp, vp, ctypes.POINTER(SolveParams), ctypes.POINTER(SolveStats)]
        L.qwsg_bicgstab_dd.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_double, ip]
        L.qwsg_norm2.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_double)]
        L.qwsg_solve_batch.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.POINTER(SolveParams), ctypes.POINTER(SolveStats)]
        L.qwsg_server_request_bytes.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.qwsg_solve_multi.argtypes = [ctypes.c_int, vp, vp, vp, vp, ctypes.POINTER(SolveParams),
                                       ctypes.POINTER(SolveStats)]
        L.qwsg_synchronize.argtypes = [vp]
        L.qwsg_time_dslash.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.qwsg_set_dslash_kernel.argtypes = [vp, ctypes.c_int]
        L.qwsg_autotune_dslash.argtypes = [vp, ctypes.c_int, ip]
        L.qwsg_set_dslash_order.argtypes = [vp, ctypes.c_int]
        L.qwsg_get_dslash_order.argtypes = [vp, ip]
        L.qwsg_autotune_solve.argtypes = [vp, vp, vp, ctypes.POINTER(SolveParams), vp, ctypes.c_int, vp, ctypes.c_int,
                                          ctypes.POINTER(SolveParams), vp]
        L.qwsg_autotune_blocks.argtypes = [vp, vp, ctypes.POINTER(SolveParams), vp, ctypes.c_int, ip, vp]
        L.qwsg_time_sap.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]
        L.qwsg_device_count.argtypes = [ip]
        L.qwsg_time_clover_rhs.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.qwsg_stream_bandwidth.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double)]
        L.qwsg_stream_bandwidth_one.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_double)]
        L.qwsg_generate_gauge.argtypes = [vp, ctypes.c_double, ctypes.c_ulonglong]
        L.qwsg_compute_clover.argtypes = [vp, ctypes.c_double]
        L.qwsg_gauge_download.argtypes = [vp, vp]
        L.qwsg_clover_download.argtypes = [vp, vp]
        L.qwsg_spinor_gaussian.argtypes = [vp, ctypes.c_int, vp, ctypes.c_ulonglong]
        L.qwsg_spinor_point.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise QwsError(lib().qwsg_last_error().decode())


def declared_functions():
    """Function names declared in include/qws_mi355x.h."""
    import re
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "for", "while", "sizeof", "return", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Spinor:
    """A device spinor field (precision F64/F32/F16) owned by the library."""

    def __init__(self, lat, prec):
        self.lat, self.prec = lat, prec
        p = ctypes.c_void_p()
        check(lib().qwsg_spinor_alloc(lat.h, prec, ctypes.byref(p)))
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr:
                lib().qwsg_spinor_free(self.ptr)
        except Exception:
            pass

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        assert a.size == self.lat.V * 24
        check(lib().qwsg_spinor_upload(self.lat.h, self.prec, self.ptr, _p(a)))
        return self

    def download(self):
        out = np.empty((self.lat.V, 3, 4, 2))
        check(lib().qwsg_spinor_download(self.lat.h, self.prec, self.ptr, _p(out)))
        return out

    def zero(self):
        check(lib().qwsg_spinor_zero(self.lat.h, self.prec, self.ptr))
        return self

    def digest(self):
        """Order-independent digest of the spinor's bytes (equal for bit-identical data)."""
        d = ctypes.c_ulonglong()
        check(lib().qwsg_spinor_digest(self.lat.h, self.prec, self.ptr, ctypes.byref(d)))
        return d.value

    def norm2(self):
        v = ctypes.c_double()
        check(lib().qwsg_norm2(self.lat.h, self.prec, self.ptr, ctypes.byref(v)))
        return v.value


class FieldSpec(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int * 4), ("B", ctypes.c_int * 4), ("eps", ctypes.c_double),
                ("seed", ctypes.c_ulonglong)]


class Request(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("kappa", ctypes.c_double), ("nrhs", ctypes.c_int),
                ("seed", ctypes.c_ulonglong), ("arrival", ctypes.c_double)]


class RequestResult(ctypes.Structure):
    _fields_ = [("arrival", ctypes.c_double), ("admit", ctypes.c_double), ("finish", ctypes.c_double),
                ("outer_iters", ctypes.c_int), ("inner_iters", ctypes.c_int), ("restarts", ctypes.c_int),
                ("true_residual_max", ctypes.c_double), ("bytes", ctypes.c_double), ("digest", ctypes.c_ulonglong)]


class Server:
    """Solve server with resident gauge fields (qwsg_server_*).

    fields: list of (L, B, eps, seed); requests: list of dicts with keys field, kappa,
    nrhs, seed, arrival (seconds).  run() replays the trace in real time and returns one
    dict per request (arrival/admit/finish times, iterations, true residual, digest)."""

    def __init__(self, fields, links32=12, tbc=-1, csw=1.0, slots=4, mem_budget=0.0, nsap=4, nm=2, tol=1e-11,
                 maxiter=40, tol_s=1e-3, maxiter_s=40, true_tol=1e-10, max_restarts=4, sap_precision=F32):
        fs = (FieldSpec * len(fields))()
        for i, (L, B, eps, seed) in enumerate(fields):
            fs[i].L[:] = list(L)
            fs[i].B[:] = list(B)
            fs[i].eps = eps
            fs[i].seed = seed
        p = SolveParams(nsap, nm, tol, maxiter, tol_s, maxiter_s, true_tol, max_restarts, sap_precision)
        self.h = ctypes.c_void_p()
        check(lib().qwsg_server_create(fs, len(fields), links32, tbc, csw, ctypes.byref(p), slots,
                                       float(mem_budget), ctypes.byref(self.h)))

    def run(self, requests):
        n = len(requests)
        rq = (Request * n)()
        for i, r in enumerate(requests):
            rq[i].field, rq[i].kappa, rq[i].nrhs = r["field"], r["kappa"], r["nrhs"]
            rq[i].seed, rq[i].arrival = r["seed"], r["arrival"]
        out = (RequestResult * n)()
        check(lib().qwsg_server_run(self.h, rq, n, out))
        return [{k: getattr(o, k) for k, _ in RequestResult._fields_} for o in out]

    def close(self):
        if self.h:
            check(lib().qwsg_server_destroy(self.h))
            self.h = None

    @staticmethod
    def request_bytes(V, nrhs):
        """Device memory the server reserves for a request of nrhs sources on V sites."""
        b = ctypes.c_double()
        check(lib().qwsg_server_request_bytes(int(V), int(nrhs), ctypes.byref(b)))
        return b.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EoSpinor:
    """A device spinor of one parity: V/2 x-compressed sites (FP64 or FP32)."""

    def __init__(self, lat, prec):
        self.lat, self.prec = lat, prec
        self.ptr = ctypes.c_void_p()
        check(lib().qwsg_eo_spinor_alloc(lat.h, prec, ctypes.byref(self.ptr)))

    def __del__(self):
        try:
            if self.ptr:
                lib().qwsg_spinor_free(self.ptr)
        except Exception:
            pass

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64).reshape(self.lat.V // 2, 24)
        check(lib().qwsg_eo_spinor_upload(self.lat.h, self.prec, self.ptr, _p(a)))

    def download(self):
        a = np.empty((self.lat.V // 2, 24))
        check(lib().qwsg_eo_spinor_download(self.lat.h, self.prec, self.ptr, _p(a)))
        return a


class Lattice:
    """One lattice with its gauge and clover fields resident on the GPU.

    L: extents (x, y, z, t); B: SAP block extents; tbc: t boundary (+1/-1);
    links32: FP32 link storage, 18 reals or 12 (third row rebuilt; SU(3) only)."""

    def __init__(self, L, B, tbc=-1, links32=18, kappa=None):
        self.L, self.B = list(L), list(B)
        self.links32 = int(links32)
        h = ctypes.c_void_p()
        Li = (ctypes.c_int * 4)(*L)
        Bi = (ctypes.c_int * 4)(*B)
        check(lib().qwsg_create(Li, Bi, int(tbc), int(links32), ctypes.byref(h)))
        self.h = h
        v = ctypes.c_longlong()
        check(lib().qwsg_volume(h, ctypes.byref(v)))
        self.V = v.value
        if kappa is not None:
            self.set_kappa(kappa)

    def close(self):
        if self.h:
            check(lib().qwsg_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sap_fused(self, on):
        check(lib().qwsg_set_sap_fused(self.h, int(bool(on))))

    def set_sap_face_buffer(self, on):
        """FP32 fused SAP face updates through the face buffer (True) or by gathers."""
        check(lib().qwsg_set_sap_face_buffer(self.h, int(bool(on))))

    def set_sap_source_update(self, on):
        """Inner BiCGStab direction update inside the first SAP sweep (True) or a separate pass."""
        check(lib().qwsg_set_sap_source_update(self.h, int(bool(on))))

    def set_kappa(self, k):
        self.kappa = float(k)
        check(lib().qwsg_set_kappa(self.h, self.kappa))

    def load_gauge(self, U):
        check(lib().qwsg_load_gauge(self.h, _p(np.ascontiguousarray(U, dtype=np.float64))))

    def load_clover(self, C):
        check(lib().qwsg_load_clover(self.h, _p(np.ascontiguousarray(C, dtype=np.float64))))

    def load_gauge_f32(self, U):
        check(lib().qwsg_load_gauge_f32(self.h, _p(np.ascontiguousarray(U, dtype=np.float32))))

    def load_clover_f32(self, C):
        check(lib().qwsg_load_clover_f32(self.h, _p(np.ascontiguousarray(C, dtype=np.float32))))

    def generate_gauge(self, eps, seed):
        check(lib().qwsg_generate_gauge(self.h, float(eps), int(seed)))

    def compute_clover(self, csw):
        check(lib().qwsg_compute_clover(self.h, float(csw)))

    def gauge(self):
        U = np.empty((4, self.V, 3, 3, 2))
        check(lib().qwsg_gauge_download(self.h, _p(U)))
        return U

    def clover_field(self):
        C = np.empty((self.V, 2, 36))
        check(lib().qwsg_clover_download(self.h, _p(C)))
        return C

    def gaussian(self, prec, seed):
        s = Spinor(self, prec)
        check(lib().qwsg_spinor_gaussian(self.h, prec, s.ptr, int(seed)))
        return s

    def point(self, prec, site, spin, colour):
        s = Spinor(self, prec)
        site = (ctypes.c_int * 4)(*site)
        check(lib().qwsg_spinor_point(self.h, prec, s.ptr, site, spin, colour))
        return s

    def spinor(self, prec, host=None):
        s = Spinor(self, prec)
        if host is not None:
            s.upload(host)
        return s

    def dslash(self, out, inp, hops=HOPS_ALL, colour=-1):
        check(lib().qwsg_dslash(self.h, inp.prec, out.ptr, inp.ptr, hops, colour))

    def face(self, out, inp, factor, colour=-1):
        check(lib().qwsg_face(self.h, inp.prec, out.ptr, inp.ptr, float(factor), colour))

    def clover(self, out, inp):
        check(lib().qwsg_clover(self.h, inp.prec, out.ptr, inp.ptr))

    def jinv(self, x, b, colour, nm):
        check(lib().qwsg_jinv(self.h, b.prec, x.ptr, b.ptr, colour, nm))

    def prec_ddd(self, out, inp, nsap, nm):
        check(lib().qwsg_prec_ddd(self.h, inp.prec, out.ptr, inp.ptr, nsap, nm))

    def prec_ddd_noalloc(self, out, inp, nsap, nm):
        """prec_ddd in FP32 through the entry point that allocates nothing."""
        check(lib().qwsg_prec_ddd_noalloc(self.h, out.ptr, inp.ptr, nsap, nm))

    def prec_s(self, out, inp, nsap, nm):
        check(lib().qwsg_prec_s(self.h, inp.prec, out.ptr, inp.ptr, nsap, nm))

    def bicgstab_precdd(self, x, b, tol, maxiter, nsap, nm):
        it = ctypes.c_int()
        check(lib().qwsg_bicgstab_precdd(self.h, b.prec, x.ptr, b.ptr, float(tol), maxiter, nsap, nm,
                                         ctypes.byref(it)))
        return it.value

    # ---- even/odd family (x-compressed parity fields of V/2 sites)
    def eo_load_gauge(self, U):
        """U: (2, 4, V/2, 18) FP64, the link from parity-p site h."""
        a = np.ascontiguousarray(U, dtype=np.float64)
        check(lib().qwsg_eo_load_gauge(self.h, _p(a)))

    def eo_load_clover(self, C):
        a = np.ascontiguousarray(C, dtype=np.float64)
        check(lib().qwsg_eo_load_clover(self.h, _p(a)))

    def eo_from_lattice(self):
        check(lib().qwsg_eo_from_lattice(self.h))

    def eo_spinor(self, prec, host=None):
        s = EoSpinor(self, prec)
        if host is not None:
            s.upload(host)
        return s

    def deo(self, out, inp, pe, set=0, dag=False, clv=False, mode=0, fac=1.0, in0=None):
        check(lib().qwsg_deo(self.h, inp.prec, pe, set, int(dag), int(clv), mode, float(fac), out.ptr, inp.ptr,
                             in0.ptr if in0 is not None else None))

    def mtilde(self, out, inp, pce=0, which=0):
        check(lib().qwsg_mtilde(self.h, inp.prec, pce, which, out.ptr, inp.ptr))

    def eo_bicgstab(self, x, b, maxiter=100, tol=1e-14, pce=0):
        """BiCGStab on mtilde (FP64 parity fields); x holds the initial guess."""
        it = ctypes.c_int()
        check(lib().qwsg_eo_bicgstab(self.h, x.ptr, b.ptr, maxiter, float(tol), pce, ctypes.byref(it)))
        return it.value

    def eo_cg(self, x, b, maxiter=100, tol=1e-14, pce=0):
        """CG on mtilde^+ mtilde (FP64 parity fields); x holds the initial guess."""
        it = ctypes.c_int()
        check(lib().qwsg_eo_cg(self.h, x.ptr, b.ptr, maxiter, float(tol), pce, ctypes.byref(it)))
        return it.value

    def eo_mcg(self, b, sigma, maxiter=100, tol=1e-14, pce=0):
        """Multi-shift CG: one FP64 parity field per shift, solving
        (mtilde^+ mtilde + sigma_k) x_k = b from zero.  Returns (fields, iterations)."""
        n = len(sigma)
        xs = [self.eo_spinor(F64) for _ in range(n)]
        xp = (ctypes.c_void_p * n)(*[x.ptr.value for x in xs])
        sg = (ctypes.c_double * n)(*[float(v) for v in sigma])
        it = ctypes.c_int()
        check(lib().qwsg_eo_mcg(self.h, xp, b.ptr, n, sg, maxiter, float(tol), pce, ctypes.byref(it)))
        return xs, it.value

    def msap_h(self, out, inp, nsap, nm):
        """out = M_SAP inp with FP16-stored SAP iterates and fields (FP32 vectors)."""
        assert out.prec == F32 and inp.prec == F32
        check(lib().qwsg_msap_h(self.h, out.ptr, inp.ptr, nsap, nm))

    def bicgstab_precdd_h(self, x, b, tol, maxiter, nsap, nm):
        assert x.prec == F32 and b.prec == F32
        it = ctypes.c_int()
        check(lib().qwsg_bicgstab_precdd_h(self.h, x.ptr, b.ptr, float(tol), maxiter, nsap, nm, ctypes.byref(it)))
        return it.value

    def bicgstab_dd(self, x, b, maxiter, tol):
        """Plain BiCGStab on D (bicgstab_dd_d.cc:64) from the initial guess in x; with
        F64 vectors the all-FP64 solve.  Returns the last iteration."""
        assert x.prec == b.prec
        it = ctypes.c_int()
        check(lib().qwsg_bicgstab_dd(self.h, x.prec, x.ptr, b.ptr, int(maxiter), float(tol), ctypes.byref(it)))
        return it.value

    def solve(self, x, b, nsap=4, nm=2, tol=-1.0, maxiter=6, tol_s=-1.0, maxiter_s=50, true_tol=-1.0,
              max_restarts=0, sap_precision=F32):
        p = SolveParams(nsap, nm, tol, maxiter, tol_s, maxiter_s, true_tol, max_restarts, sap_precision)
        st = SolveStats()
        check(lib().qwsg_solve(self.h, x.ptr, b.ptr, ctypes.byref(p), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in SolveStats._fields_}

    def solve_batch(self, xs, bs, nsap=4, nm=2, tol=-1.0, maxiter=6, tol_s=-1.0, maxiter_s=50, true_tol=-1.0,
                    max_restarts=0, sap_precision=F32):
        n = len(xs)
        assert len(bs) == n
        p = SolveParams(nsap, nm, tol, maxiter, tol_s, maxiter_s, true_tol, max_restarts, sap_precision)
        st = (SolveStats * n)()
        xp = (ctypes.c_void_p * n)(*[x.ptr.value for x in xs])
        bp = (ctypes.c_void_p * n)(*[b.ptr.value for b in bs])
        check(lib().qwsg_solve_batch(self.h, n, xp, bp, ctypes.byref(p), st))
        return [{k: getattr(s, k) for k, _ in SolveStats._fields_} for s in st]

    def sync(self):
        check(lib().qwsg_synchronize(self.h))

    def time_dslash(self, prec, reps):
        ms = ctypes.c_double()
        check(lib().qwsg_time_dslash(self.h, prec, reps, ctypes.byref(ms)))
        return ms.value

    def set_dslash_kernel(self, k):
        """FP32 full-operator kernel on 256-site blocks: 0 global gathers, 3/4 LDS-staged."""
        check(lib().qwsg_set_dslash_kernel(self.h, int(k)))

    def autotune_solve(self, x, b, nsaps, nms, **params):
        """Time one solve of b per (nsap, nm) candidate pair; returns (best params dict,
        seconds[len(nsaps)][len(nms)], negative where the true residual was missed)."""
        base = dict(dict(nsap=4, nm=2, tol=-1.0, maxiter=6, tol_s=-1.0, maxiter_s=50, true_tol=-1.0, max_restarts=0,
                         sap_precision=F32), **params)
        p = SolveParams(base["nsap"], base["nm"], base["tol"], base["maxiter"], base["tol_s"], base["maxiter_s"],
                        base["true_tol"], base["max_restarts"], base["sap_precision"])
        best = SolveParams()
        ns = (ctypes.c_int * len(nsaps))(*nsaps)
        nm = (ctypes.c_int * len(nms))(*nms)
        sec = (ctypes.c_double * (len(nsaps) * len(nms)))()
        check(lib().qwsg_autotune_solve(self.h, x.ptr, b.ptr, ctypes.byref(p), ns, len(nsaps), nm, len(nms),
                                        ctypes.byref(best), sec))
        out = dict(base, nsap=best.nsap, nm=best.nm)
        return out, [[sec[i * len(nms) + j] for j in range(len(nms))] for i in range(len(nsaps))]

    def autotune_blocks(self, b, shapes, **params):
        """Time one solve of b on a copy of this lattice for each candidate block shape;
        returns (fastest shape, seconds per shape, negative where the shape does not tile
        the lattice or missed the true residual).  This lattice keeps its blocks."""
        base = dict(dict(nsap=4, nm=2, tol=-1.0, maxiter=6, tol_s=-1.0, maxiter_s=50, true_tol=-1.0, max_restarts=0,
                         sap_precision=F32), **params)
        p = SolveParams(base["nsap"], base["nm"], base["tol"], base["maxiter"], base["tol_s"], base["maxiter_s"],
                        base["true_tol"], base["max_restarts"], base["sap_precision"])
        n = len(shapes)
        sh = (ctypes.c_int * (4 * n))(*[int(v) for s in shapes for v in s])
        sec = (ctypes.c_double * n)()
        best = ctypes.c_int()
        check(lib().qwsg_autotune_blocks(self.h, b.ptr, ctypes.byref(p), sh, n, ctypes.byref(best), sec))
        return list(shapes[best.value]), [sec[i] for i in range(n)]

    def set_dslash_order(self, order):
        """Block order of the launches over every block: 0 index, 1 t-z pencils per XCD."""
        check(lib().qwsg_set_dslash_order(self.h, int(order)))

    def dslash_order(self):
        o = ctypes.c_int()
        check(lib().qwsg_get_dslash_order(self.h, ctypes.byref(o)))
        return o.value

    def autotune_dslash(self, reps=20):
        """Time the full-operator kernels on this lattice, keep and return the fastest."""
        k = ctypes.c_int()
        check(lib().qwsg_autotune_dslash(self.h, int(reps), ctypes.byref(k)))
        return k.value

    def time_sap(self, prec, reps, nsap, nm):
        """Average device ms of one fused SAP launch (FP32 or FP16 storage)."""
        ms = ctypes.c_double()
        check(lib().qwsg_time_sap(self.h, prec, reps, nsap, nm, ctypes.byref(ms)))
        return ms.value


def stream_bandwidth(nbytes, nin=1, reps=20, one_stream=False):
    """Measured GB/s of a streaming kernel summing nin arrays of nbytes into one
    (one_stream: the nin inputs are consecutive chunks of one array, read region by
    region)."""
    g = ctypes.c_double()
    f = lib().qwsg_stream_bandwidth_one if one_stream else lib().qwsg_stream_bandwidth
    check(f(int(nbytes), int(nin), int(reps), ctypes.byref(g)))
    return g.value


def device_count():
    n = ctypes.c_int()
    check(lib().qwsg_device_count(ctypes.byref(n)))
    return n.value


def solve_multi(parts, nsap=4, nm=2, tol=-1.0, maxiter=6, tol_s=-1.0, maxiter_s=50, true_tol=-1.0,
                max_restarts=0, sap_precision=F32):
    """qwsg_solve_multi: parts = [(lattice, [x...], [b...]), ...]; the lattices solve
    their right-hand sides at once, the SAP launches fused over them.  Returns the
    stats per lattice and right-hand side."""
    n = len(parts)
    lats = (ctypes.c_void_p * n)(*[lat.h for lat, _, _ in parts])
    nr = (ctypes.c_int * n)(*[len(xs) for _, xs, _ in parts])
    xs = [x for _, xl, _ in parts for x in xl]
    bs = [b for _, _, bl in parts for b in bl]
    assert len(xs) == len(bs)
    tot = len(xs)
    p = SolveParams(nsap, nm, tol, maxiter, tol_s, maxiter_s, true_tol, max_restarts, sap_precision)
    st = (SolveStats * tot)()
    xp = (ctypes.c_void_p * tot)(*[x.ptr.value for x in xs])
    bp = (ctypes.c_void_p * tot)(*[b.ptr.value for b in bs])
    check(lib().qwsg_solve_multi(n, lats, nr, xp, bp, ctypes.byref(p), st))
    out, at = [], 0
    for _, xl, _ in parts:
        out.append([{k: getattr(s, k) for k, _ in SolveStats._fields_} for s in st[at:at + len(xl)]])
        at += len(xl)
    return out


def time_clover_rhs(V, nrep, reps):
    """The multi-RHS clover product (16 RHS per site, FP16 storage, FP32 accumulation)
    on the vector ALUs and on the matrix cores: (ms VALU, ms MFMA, max relative
    difference).  Measurement only."""
    ms = (ctypes.c_double * 2)()
    d = ctypes.c_double()
    check(lib().qwsg_time_clover_rhs(int(V), int(nrep), int(reps), ms, ctypes.byref(d)))
    return ms[0], ms[1], d.value
