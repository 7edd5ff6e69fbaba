"""Run the QWS reference library (oracle/_ref/libqws_ref.so) on given inputs.

TEST INFRASTRUCTURE ONLY.  The reference keeps one lattice geometry per process
(static work arrays sized at first use, qws.cc:776-790, bicgstab_precdd_s.cc:85-115),
so every geometry runs in its own child process:

    python -m oracle.ref_worker IN.npz OUT.npz

IN.npz holds nloc (lx,ly,lz,lt), kappa, U, C (plain layout, FP64), b, x0, the
solver parameters and the list of operations; OUT.npz holds the results in the
plain layout.  Single rank only (no MPI in this build of the reference), so the
fbc is (1,1,1,-1) as in the reference driver (main.cc:154-158).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFLIB = os.path.join(HERE, "_ref", "libqws_ref.so")
VLEND, VLENS = 8, 16  # the layout this build of the reference uses (oracle/Makefile.ref)


def available():
    return os.path.exists(REFLIB)


# ---------------------------------------------------------------- layout maps
# Host AoSoA of the reference (qws.h:111-192; index maps main.cc:426-429,485-488):
# site (x,y,z,t) -> domain d = x // nxh, vector i = (x%nxh)//VLEN + nxv*(y + ny*(z + nz*t)),
# lane j = (x%nxh) % VLEN.

def spinor_to_aosoa(a, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    p = np.asarray(a).reshape(nt, nz, ny, 2, nxv, vlen, 3, 4, 2)
    return np.ascontiguousarray(p.transpose(3, 0, 1, 2, 4, 6, 7, 8, 5))


def spinor_from_aosoa(a, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    p = np.asarray(a).reshape(2, nt, nz, ny, nxv, 3, 4, 2, vlen)
    return np.ascontiguousarray(p.transpose(1, 2, 3, 0, 4, 8, 5, 6, 7)).reshape(-1, 3, 4, 2)


def links_to_aosoa(U, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    p = np.asarray(U).reshape(4, nt, nz, ny, 2, nxv, vlen, 3, 3, 2)
    return np.ascontiguousarray(p.transpose(4, 0, 1, 2, 3, 5, 7, 8, 9, 6))


def clover_to_aosoa(C, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    p = np.asarray(C).reshape(nt, nz, ny, 2, nxv, vlen, 2, 36)
    return np.ascontiguousarray(p.transpose(3, 0, 1, 2, 4, 6, 7, 5))


# Even/odd reading of the same host arrays (deo_in_d.cc:65-90, qws_bqcd.cc:82-85):
# a field of one parity p holds vold vectors; vector i = xv + nxv*(y + ny*(z + nz*t)),
# lane j, is the site with compressed x index xh = xv*vlen + j.  Links:
# glud[i + vold*mu + 4*vold*p]; clover clvd[i + vold*p].  The even/odd oracle works on
# x-compressed plain fields: spinor (V/2, 24) with h = xh + (Lx/2)*(y + Ly*(z + Lz*t)),
# links (2, 4, V/2, 18), clover (2, V/2, 72).

def eo_spinor_from_raw(raw, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    r = np.asarray(raw).reshape(nt, nz, ny, nxv, 3, 4, 2, vlen).transpose(0, 1, 2, 3, 7, 4, 5, 6)
    return np.ascontiguousarray(r).reshape(-1, 24)


def eo_spinor_to_raw(a, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    r = np.asarray(a).reshape(nt, nz, ny, nxv, vlen, 3, 4, 2).transpose(0, 1, 2, 3, 5, 6, 7, 4)
    return np.ascontiguousarray(r)


def eo_links_from_raw(raw, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    r = np.asarray(raw).reshape(2, 4, nt, nz, ny, nxv, 3, 3, 2, vlen).transpose(0, 1, 2, 3, 4, 5, 9, 6, 7, 8)
    return np.ascontiguousarray(r).reshape(2, 4, -1, 18)


def eo_clover_from_raw(raw, nloc, vlen):
    nx, ny, nz, nt = nloc
    nxv = nx // 2 // vlen
    r = np.asarray(raw).reshape(2, nt, nz, ny, nxv, 2, 36, vlen).transpose(0, 1, 2, 3, 4, 7, 5, 6)
    return np.ascontiguousarray(r).reshape(2, -1, 72)


class Ref:
    def __init__(self, nloc, kappa):
        self.lib = ctypes.CDLL(REFLIB)
        self.nloc = list(nloc)
        lx, ly, lz, lt = [ctypes.c_int(v) for v in nloc]
        self._keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1),
                      ctypes.c_int(0), ctypes.c_int(1), (ctypes.c_int * 4)(1, 2, 2, 2)]
        npe, fbc, pce, pco, bs = self._keep
        self.lib.qws_init_(ctypes.byref(lx), ctypes.byref(ly), ctypes.byref(lz), ctypes.byref(lt),
                           npe, fbc, ctypes.byref(pce), ctypes.byref(pco), bs)
        ctypes.c_double.in_dll(self.lib, "kappa").value = kappa
        ctypes.c_double.in_dll(self.lib, "mkappa").value = -kappa
        ctypes.c_double.in_dll(self.lib, "kappa2").value = kappa * kappa
        self.vold = ctypes.c_int.in_dll(self.lib, "vold").value
        self.vols = ctypes.c_int.in_dll(self.lib, "vols").value

    def _global(self, name, ctype, n):
        addr = ctypes.c_void_p.in_dll(self.lib, name).value
        return np.ctypeslib.as_array((ctype * n).from_address(addr))

    def set_fields(self, U, C):
        n = self.nloc
        gd = links_to_aosoa(U, n, VLEND).ravel()
        gs = links_to_aosoa(U.astype(np.float32), n, VLENS).ravel()
        cd = clover_to_aosoa(C, n, VLEND).ravel()
        cs = clover_to_aosoa(C.astype(np.float32), n, VLENS).ravel()
        self._global("glud", ctypes.c_double, gd.size)[:] = gd
        self._global("glus", ctypes.c_float, gs.size)[:] = gs
        self._global("clvd", ctypes.c_double, cd.size)[:] = cd
        self._global("clvs", ctypes.c_float, cs.size)[:] = cs

    def d(self, a):
        return spinor_to_aosoa(a, self.nloc, VLEND)

    def s(self, a):
        return spinor_to_aosoa(a.astype(np.float32), self.nloc, VLENS)

    def back_d(self, a):
        return spinor_from_aosoa(a, self.nloc, VLEND)

    def back_s(self, a):
        return spinor_from_aosoa(a, self.nloc, VLENS)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i(v):
    return ctypes.byref(ctypes.c_int(v))


def _dbl(v):
    return ctypes.byref(ctypes.c_double(v))


def run(inp):
    nloc = [int(v) for v in inp["nloc"]]
    kappa = float(inp["kappa"])
    r = Ref(nloc, kappa)
    r.set_fields(inp["U"], inp["C"])
    return run_ops(r, inp)


def run_ops(r, inp):
    """The operations inp["ops"] on the reference instance r (fields already set)."""
    lib = r.lib
    kappa = ctypes.c_double.in_dll(lib, "kappa").value
    vols, vold = r.vols, r.vold
    b = inp["b"]
    nsap, nm = int(inp["nsap"]), int(inp["nm"])
    res = {}
    for op in [str(o) for o in inp["ops"]]:
        if op == "ddd_d":
            out = np.zeros_like(r.d(b))
            lib.ddd_d_(_ptr(out), _ptr(r.d(b)))
            res[op] = r.back_d(out)
        elif op == "ddd_s":
            out = np.zeros_like(r.s(b))
            lib.ddd_s_(_ptr(out), _ptr(r.s(b)))
            res[op] = r.back_s(out)
        elif op.startswith("ddd_in_s"):  # ddd_in_s0 / ddd_in_s1: domain d
            d = int(op[-1])
            bs = r.s(b)
            out = np.zeros_like(bs)
            lib.ddd_in_s_(_ptr(out[d]), _ptr(bs[d]), _i(d))
            res[op] = r.back_s(out)
        elif op.startswith("ddd_out_s"):  # face term for domain d with factor -kappa
            d = int(op[-1])
            bs = r.s(b)
            out = np.zeros_like(bs)
            lib.ddd_out_pos_s_.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]
            lib.ddd_out_pre_s_(_ptr(bs), _i(d))
            lib.ddd_out_pos_s_(_ptr(out[d]), _ptr(bs), _i(d), ctypes.c_float(-kappa))
            res[op] = r.back_s(out)
        elif op.startswith("jinv"):  # jinv0 / jinv1
            d = int(op[-1])
            bs = r.s(b)
            out = np.zeros_like(bs)
            lib.jinv_ddd_in_s_(_ptr(out[d]), _ptr(bs[d]), _i(d), _i(nm))
            res[op] = r.back_s(out)
        elif op == "prec_ddd_s":
            bs = r.s(b)
            out = np.zeros_like(bs)
            lib.prec_ddd_s_(_ptr(out), _ptr(bs), _i(nsap), _i(nm))
            res[op] = r.back_s(out)
        elif op == "prec_s":
            bs = r.s(b)
            out = np.zeros_like(bs)
            lib.prec_s_(_ptr(out), _ptr(bs), _i(nsap), _i(nm))
            res[op] = r.back_s(out)
        elif op == "bicgstab_precdd_s":
            bs = r.s(b)
            out = np.zeros_like(bs)
            conv = ctypes.c_int(0)
            lib.bicgstab_precdd_s_(_ptr(out), _ptr(bs), _dbl(float(inp["tol_s"])), ctypes.byref(conv),
                                   _i(int(inp["maxiter_s"])), _i(nsap), _i(nm))
            res[op] = r.back_s(out)
        elif op == "time_inner":  # wall time of bicgstab_precdd_s_ with maxiter_s fixed iterations
            import time
            bs = r.s(b / np.sqrt(np.sum(np.asarray(b, dtype=np.float64) ** 2)))
            out = np.zeros_like(bs)
            conv = ctypes.c_int(0)
            t0 = time.perf_counter()
            lib.bicgstab_precdd_s_(_ptr(out), _ptr(bs), _dbl(-1.0), ctypes.byref(conv),
                                   _i(int(inp["maxiter_s"])), _i(nsap), _i(nm))
            res[op] = np.array(time.perf_counter() - t0)
        elif op == "bicgstab_dd_mix":
            x = r.d(inp["x0"])
            conv = ctypes.c_int(0)
            lib.bicgstab_dd_mix_(_ptr(x), _ptr(r.d(b)), _dbl(float(inp["tol"])), ctypes.byref(conv),
                                 _i(int(inp["maxiter"])), _dbl(float(inp["tol_s"])), _i(int(inp["maxiter_s"])),
                                 _i(nsap), _i(nm))
            res[op] = r.back_d(x)
            res[op + "_iter"] = np.array(conv.value)
        elif op in ("bicgstab_dd_s", "bicgstab_dd_d"):  # plain BiCGStab on D from x0
            fp64 = op.endswith("_d")
            x = r.d(inp["x0"]) if fp64 else r.s(inp["x0"])
            bb = r.d(b) if fp64 else r.s(b)
            conv = ctypes.c_int(0)
            getattr(lib, op + "_")(_ptr(x), _ptr(bb), ctypes.byref(conv), _i(int(inp["maxiter"])))
            res[op] = r.back_d(x) if fp64 else r.back_s(x)
            res[op + "_iter"] = np.array(conv.value)
        elif op in ("cg_vm", "mcg_vm"):
            # even/odd solvers on mtdagmt_vm_: the parity field is the first x-domain of the
            # driver's AoSoA spinor (as main.cc:200-213 passes it); also runs the oracle's
            # restatement (oracle/eo_solvers.py) with the reference's own operator
            from oracle import eo_solvers as ES
            bd = np.ascontiguousarray(r.d(b).reshape(2 * vold, -1)[:vold])
            mi = int(inp["maxiter"])

            def apply(v):
                out = np.zeros_like(bd)
                lib.mtdagmt_vm_(_ptr(out), _ptr(np.ascontiguousarray(v.reshape(bd.shape))))
                return out.ravel()

            if op == "cg_vm":
                x = np.zeros_like(bd)
                conv = ctypes.c_int(0)
                lib.cg_vm_(_ptr(x), _ptr(bd), ctypes.byref(conv), _i(mi))
                res[op] = x
                xo, it = ES.cg(apply, bd.ravel(), np.zeros(bd.size), mi)
                res[op + "_oracle"] = xo.reshape(bd.shape)
                res[op + "_oracle_iter"] = np.array(it)
            else:
                sg = np.ascontiguousarray(inp["sigma"], dtype=np.float64)
                n = len(sg)
                xx = np.zeros((n,) + bd.shape)
                lib.mcg_vm_(_ptr(xx), _ptr(bd), _i(n), sg.ctypes.data_as(ctypes.c_void_p), _i(mi))
                res[op] = xx
                xs, it = ES.mcg(apply, bd.ravel(), list(sg), mi)
                res[op + "_oracle"] = np.stack([v.reshape(bd.shape) for v in xs])
                res[op + "_oracle_iter"] = np.array(it)
            res[op + "_b"] = bd
        else:
            raise ValueError(op)
    return res


def main():
    inp = dict(np.load(sys.argv[1], allow_pickle=False))
    res = run(inp)
    np.savez(sys.argv[2], **res)


if __name__ == "__main__":
    main()


def call(nloc, kappa, U, C, b, ops, x0=None, nsap=4, nm=2, tol=-1.0, maxiter=6, tol_s=-1.0,
         maxiter_s=50, timeout=600, sigma=()):
    """Run `ops` in a child process and return {op: plain-layout array}."""
    import subprocess
    import tempfile
    root = os.path.dirname(HERE)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.npz"), os.path.join(td, "out.npz")
        np.savez(fin, nloc=np.array(nloc), kappa=np.array(kappa), U=U, C=C, b=b,
                 x0=x0 if x0 is not None else np.zeros_like(b), ops=np.array(ops), nsap=np.array(nsap),
                 nm=np.array(nm), tol=np.array(tol), maxiter=np.array(maxiter), tol_s=np.array(tol_s),
                 maxiter_s=np.array(maxiter_s), sigma=np.array(sigma, dtype=np.float64))
        p = subprocess.run([sys.executable, "-m", "oracle.ref_worker", fin, fout], cwd=root,
                           capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError("reference worker failed:\n" + p.stdout[-2000:] + p.stderr[-2000:])
        return dict(np.load(fout, allow_pickle=False))
