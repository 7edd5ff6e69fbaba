"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` use this module as the reference results to compare
with.  The product
package ``qws_amd`` never imports it.

Arrays are numpy, in the oracle's plain site-major layout:
spinor (V, 3, 4, 2), links (4, V, 3, 3, 2), clover inverse (V, 2, 36); the site
index is x + Lx*(y + Ly*(z + Lz*t)).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

HOPS_ALL, HOPS_INSIDE, HOPS_FACE = 0, 1, 2


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        vp, dp, fp, ip = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ip, ip, ctypes.c_int, ctypes.c_double]
        L.orc_destroy.argtypes = [vp]
        L.orc_volume.restype = ctypes.c_long
        L.orc_volume.argtypes = [vp]
        L.orc_improper.argtypes = [vp]
        L.orc_set_fields.argtypes = [vp, dp, dp]
        L.orc_set_fields_s.argtypes = [vp, fp, fp]
        for n in ("orc_dslash_d", "orc_dslash_s"):
            getattr(L, n).argtypes = [vp, dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_face_d.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_face_s.argtypes = [vp, fp, fp, ctypes.c_float, ctypes.c_int]
        L.orc_face_qws_uniform_s.argtypes = [vp, ip, ip, ctypes.c_int, ctypes.c_float, fp, fp]
        L.orc_clover_d.argtypes = [vp, dp, dp]
        L.orc_jinv_s.argtypes = [vp, fp, fp, ctypes.c_int, ctypes.c_int]
        L.orc_prec_ddd_s.argtypes = [vp, fp, fp, ctypes.c_int, ctypes.c_int]
        L.orc_prec_s.argtypes = [vp, fp, fp, ctypes.c_int, ctypes.c_int]
        L.orc_msap_h.argtypes = [vp, fp, fp, ctypes.c_int, ctypes.c_int]
        L.orc_to_half.argtypes = [ctypes.c_long, fp, fp]
        L.orc_deo_d.argtypes = [vp, dp, dp] + [ctypes.c_int] * 5 + [ctypes.c_double, dp, dp, dp]
        L.orc_deo_s.argtypes = [vp, fp, fp] + [ctypes.c_int] * 5 + [ctypes.c_float, fp, fp, fp]
        L.orc_prec_ddd_d.argtypes = [vp, dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_bicgstab_precdd_s.argtypes = [vp, fp, fp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_bicgstab_dd_mix.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
        L.orc_norm2_d.restype = ctypes.c_double
        L.orc_norm2_d.argtypes = [ctypes.c_long, dp]
        L.orc_norm2_s.restype = ctypes.c_double
        L.orc_norm2_s.argtypes = [ctypes.c_long, fp]
        L.orc_case_inputs.argtypes = [ip, ip, dp, dp, dp, dp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _ints(v):
    return (ctypes.c_int * len(v))(*v)


def case_inputs(nloc, npe=(1, 1, 1, 1)):
    """The reference driver's inputs (main.cc:406-551) tiled over a rank grid.

    Returns (L, U, C, b, x0) with L the global extents."""
    L = [nloc[m] * npe[m] for m in range(4)]
    V = L[0] * L[1] * L[2] * L[3]
    U = np.zeros((4, V, 3, 3, 2))
    C = np.zeros((V, 2, 36))
    b = np.zeros((V, 3, 4, 2))
    x0 = np.zeros((V, 3, 4, 2))
    lib().orc_case_inputs(_ints(nloc), _ints(npe), _p(U), _p(C), _p(b), _p(x0))
    return L, U, C, b, x0


class Oracle:
    """One lattice geometry with its fields.  B = block (SAP domain) extents."""

    def __init__(self, L, B, tbc, kappa, U, C, Us=None, Cs=None):
        self.L, self.B = list(L), list(B)
        self.kappa = kappa
        self._keep = [np.ascontiguousarray(U, dtype=np.float64), np.ascontiguousarray(C, dtype=np.float64)]
        self.h = lib().orc_create(_ints(L), _ints(B), int(tbc), float(kappa))
        self.V = lib().orc_volume(self.h)
        lib().orc_set_fields(self.h, _p(self._keep[0]), _p(self._keep[1]))
        if Us is not None:
            self._keep += [np.ascontiguousarray(Us, dtype=np.float32), np.ascontiguousarray(Cs, dtype=np.float32)]
            lib().orc_set_fields_s(self.h, _p(self._keep[2]), _p(self._keep[3]))

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    @property
    def improper(self):
        return bool(lib().orc_improper(self.h))

    def _new(self, dtype, like=None):
        return np.zeros((self.V, 3, 4, 2), dtype=dtype) if like is None else np.array(like, dtype=dtype, copy=True)

    def dslash(self, psi, hops=HOPS_ALL, colour=-1, out=None):
        """out = psi - kappa T^-1 H_hops psi on sites of `colour` (others: out unchanged)."""
        dt = psi.dtype
        psi = np.ascontiguousarray(psi)
        out = self._new(dt) if out is None else out
        f = lib().orc_dslash_d if dt == np.float64 else lib().orc_dslash_s
        f(self.h, _p(out), _p(psi), hops, colour)
        return out

    def face(self, out, psi, factor, colour):
        psi = np.ascontiguousarray(psi)
        if psi.dtype == np.float64:
            lib().orc_face_d(self.h, _p(out), _p(psi), float(factor), colour)
        else:
            lib().orc_face_s(self.h, _p(out), _p(psi), float(factor), colour)
        return out

    def face_qws_uniform_s(self, out, psi, nloc, npe, d, factor):
        lib().orc_face_qws_uniform_s(self.h, _ints(nloc), _ints(npe), d, float(factor), _p(out),
                                     _p(np.ascontiguousarray(psi)))
        return out

    def clover(self, psi):
        out = self._new(np.float64)
        lib().orc_clover_d(self.h, _p(out), _p(np.ascontiguousarray(psi, dtype=np.float64)))
        return out

    def jinv_s(self, b, colour, nm, x=None):
        x = self._new(np.float32) if x is None else x
        lib().orc_jinv_s(self.h, _p(x), _p(np.ascontiguousarray(b, dtype=np.float32)), colour, nm)
        return x

    def prec_ddd_s(self, psi, nsap, nm):
        out = self._new(np.float32)
        lib().orc_prec_ddd_s(self.h, _p(out), _p(np.ascontiguousarray(psi, dtype=np.float32)), nsap, nm)
        return out

    def prec_ddd_d(self, psi, nsap, nm):
        out = self._new(np.float64)
        lib().orc_prec_ddd_d(self.h, _p(out), _p(np.ascontiguousarray(psi, dtype=np.float64)), nsap, nm)
        return out

    def prec_s(self, psi, nsap, nm):
        out = self._new(np.float32)
        lib().orc_prec_s(self.h, _p(out), _p(np.ascontiguousarray(psi, dtype=np.float32)), nsap, nm)
        return out

    def msap_h(self, psi, nsap, nm):
        """M_SAP with FP16 storage (assign_mult_mSAP_s_h_: nsap - 1 sweeps)."""
        out = self._new(np.float32)
        lib().orc_msap_h(self.h, _p(out), _p(np.ascontiguousarray(psi, dtype=np.float32)), nsap, nm)
        return out

    def deo(self, Ueo, Ceo, pe, psi, set=0, dag=False, clv=False, mode=0, fac=1.0, out=None, in0=None):
        """Even/odd hopping (qws_oracle.cpp deo) on x-compressed parity fields of V/2 sites:
        Ueo (2,4,V/2,18), Ceo (2,V/2,72); psi the parity-(1-pe) field, the result the
        parity-pe field.  set 0 all hops, 1 hops that do not wrap, 2 hops that wrap;
        mode 0 out = acc, 1 out += fac acc, 2 out = in0 + fac acc.  FP64 or FP32 after
        the dtype of psi."""
        dt = np.float64 if np.asarray(psi).dtype == np.float64 else np.float32
        vh = self.V // 2
        psi = np.ascontiguousarray(psi, dtype=dt).reshape(vh, 24)
        Ueo = np.ascontiguousarray(Ueo, dtype=dt)
        Ceo = np.ascontiguousarray(Ceo, dtype=dt)
        assert Ueo.size == 2 * 4 * vh * 18 and Ceo.size == 2 * vh * 72
        out = np.zeros((vh, 24), dtype=dt) if out is None else out
        if out.dtype != dt or not out.flags.c_contiguous or out.size != vh * 24:
            raise ValueError("out must be a contiguous parity field of the input's dtype")
        in0 = psi if in0 is None else np.ascontiguousarray(in0, dtype=dt).reshape(vh, 24)
        f = lib().orc_deo_d if dt == np.float64 else lib().orc_deo_s
        f(self.h, _p(Ueo), _p(Ceo), pe, set, int(dag), int(clv), mode, fac, _p(out), _p(psi), _p(in0))
        return out

    def bicgstab_precdd_s(self, b, tol, maxiter, nsap, nm):
        x = self._new(np.float32)
        it = lib().orc_bicgstab_precdd_s(self.h, _p(x), _p(np.ascontiguousarray(b, dtype=np.float32)),
                                        float(tol), maxiter, nsap, nm)
        return x, it

    def bicgstab_dd_mix(self, x0, b, tol, maxiter, tol_s, maxiter_s, nsap, nm):
        x = np.array(x0, dtype=np.float64, copy=True)
        inner = ctypes.c_int(0)
        it = lib().orc_bicgstab_dd_mix(self.h, _p(x), _p(np.ascontiguousarray(b, dtype=np.float64)),
                                      float(tol), maxiter, float(tol_s), maxiter_s, nsap, nm,
                                      ctypes.byref(inner))
        return x, it, inner.value


def to_half(a):
    """binary16 round-to-nearest-even of float32 values, returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty_like(a)
    lib().orc_to_half(a.size, _p(a), _p(out))
    return out


def norm2(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return lib().orc_norm2_d(a.size, _p(a))
    return lib().orc_norm2_s(a.size, _p(a))
