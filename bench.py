"""Benchmark: plain mixed-precision SAP-BiCGStab solve on one MI355X.

Workload (one step = one solve from x = 0 to an FP64 true residual
|b - D x| / |b| <= 1e-10): 32^3 x 64 lattice, SU(3) links U = exp(0.3 i H) (Philox
seed 2), kappa = 0.1365, c_sw = 1.0 clover term built from the links, antiperiodic
t, Gaussian FP64 source (seed 2); FP64 outer BiCGStab with FP32 SAP-preconditioned
inner BiCGStab on 4^4 blocks; FP32 links stored as 12 reals.

Prints one JSON line.  value = solves/s; roofline = the dominant kernel of the
solve, the fused FP32 SAP half-sweep (about three quarters of the device time),
timed with HIP events on the library's stream, its algorithmic FP32 flops against
the MI355X FP32 peak; roofline_d = the FP32 dslash kernel (full D = 1 - kappa
T^-1 H) against the HBM peak; cpu_baseline = the QWS reference library
(oracle/_ref) timed on this host on a bounded sample of the same solve.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 (vector, with packed FP32 FMA)
WORKLOAD = dict(L=[32, 32, 32, 64], B=[4, 4, 4, 4], eps=0.3, seed=2, kappa=0.1365, csw=1.0, tbc=-1, links32=12)
# Solver settings of the measured workload (chosen from a sweep of nsap, nm and tol_s on MI355X)
SOLVE = dict(nsap=2, nm=6, tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=80, true_tol=1e-10, max_restarts=4)
# algorithmic bytes per site of one FP32 D application with 12-real links:
# 4 links x 12 x 4 B + clover 72 x 4 B + input 24 x 4 B + output 24 x 4 B
BYTES_PER_SITE_F32_12 = 4 * 12 * 4 + 72 * 4 + 24 * 4 + 24 * 4


def sap_flops_per_site(B, nsap, nm):
    """Algorithmic flops per site of one fused SAP launch (the sites of one colour),
    averaged over the 2*nsap+2 launches of one preconditioner application.

    One hop along x, y or z: spin projection 12 + SU(3) times two colour vectors 132
    + accumulation of the 4-spinor 24 = 168; along t (diagonal gamma_4, so the
    projection is a selection and only two spin components are accumulated)
    132 + 12 = 144.  Hops inside a block per site: 2 (B_mu - 1) / B_mu along mu;
    across a block face: 2 / B_mu.  Clover inverse: two packed Hermitian 6x6
    products 2 x 252 + spin rotation in and out 48 = 552.  Jacobi step x = s +
    kappa A x: in-block hops + clover + 48.  Launch sequence (qws.cc:809-925): the
    first sweep's first colour has no face prologue; every other launch starts with
    the face update (face hops + clover + 48); the epilogues are the residual
    update (in-block hops + clover + 96) for the 2*nsap sweep launches, then D_in
    (in-block hops + clover + 48) and D_in plus the face term (+ 96).
    """
    hop = [168.0, 168.0, 168.0, 144.0]
    h_in = sum(2.0 * (b - 1) / b * hop[mu] for mu, b in enumerate(B))
    h_face = sum(2.0 / b * hop[mu] for mu, b in enumerate(B))
    clv = 552.0
    jac = nm * (h_in + clv + 48.0)
    pro = h_face + clv + 48.0
    e_s, e_o, e_of = h_in + clv + 96.0, h_in + clv + 48.0, h_in + clv + 96.0
    n = 2 * nsap + 2
    total = n * jac + 2 * nsap * e_s + e_o + e_of + (n - 1) * pro
    return total / n


# Compulsory HBM bytes per site of the lattice for one FP32 inner BiCGStab iteration
# (fields read once, outputs written once): two preconditioner applications, each
# 2 nsap + 2 fused SAP launches over V/2 sites at 960 B (links 192, clover 288, source or
# iterate 96, the other colour's face data 96, the residual update's source 96, two
# stored spinors 192) and one final face-term launch over V/2 sites at 624 B (clover 288,
# output read and written 192, face slots 96, the backward hops' own links 48); the
# vector updates (r only: 3 spinor passes; x and r: 7) and the dot products (4 passes)
# at 96 B per spinor pass.
def inner_iteration_bytes_per_site(nsap):
    sap = 2 * (2 * nsap + 2) * 0.5 * 960
    face = 2 * 0.5 * 624
    vectors = (3 + 7 + 4) * 96
    return sap + face + vectors


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def setup(Q, w=WORKLOAD):
    lat = Q.Lattice(w["L"], w["B"], w["tbc"], w["links32"], kappa=w["kappa"])
    lat.generate_gauge(w["eps"], w["seed"])
    lat.compute_clover(w["csw"])
    b = lat.gaussian(Q.F64, w["seed"])
    return lat, b


def run_solves(Q, lat, b, n, params):
    x = lat.spinor(Q.F64)
    stats = []
    lat.sync()
    t0 = time.perf_counter()
    for _ in range(n):
        x.zero()
        stats.append(lat.solve(x, b, **params))
    lat.sync()
    return time.perf_counter() - t0, stats, x


def fp64_reference(Q, lat, b, tol=1e-12, maxiter=20000):
    """The all-FP64 solve a mixed-precision result is compared with: plain BiCGStab on
    D (bicgstab_dd_d.cc:64) with FP64 vectors and reductions from x = 0, no preconditioner
    and no FP32 anywhere.  Returns the solution and its FP64 true residual."""
    x = lat.spinor(Q.F64)
    it = lat.bicgstab_dd(x, b, maxiter, tol)
    r = lat.spinor(Q.F64)
    lat.dslash(r, x)
    rr, bb = r.download(), b.download()
    res = float(np.sqrt(np.sum((bb - rr) ** 2) / np.sum(bb ** 2)))
    return x, res, it


def error_vs_fp64(Q, lat, x, b):
    """|x - x64| / |x64| against the all-FP64 solve of the same source."""
    xr, res, it = fp64_reference(Q, lat, b)
    a, r = x.download(), xr.download()
    return {"error": float(np.linalg.norm(a - r) / np.linalg.norm(r)), "fp64_true_residual": res,
            "fp64_iterations": it}


def batched_point_sources(Q, lat, params):
    """The 12 spin-colour point sources at the origin solved as one batch."""
    bs = [lat.point(Q.F64, (0, 0, 0, 0), sp, c) for c in range(3) for sp in range(4)]
    xs = [lat.spinor(Q.F64) for _ in bs]
    # warm-up: one outer iteration allocates the batch's work buffers
    lat.solve_batch(xs, bs, **dict(params, maxiter=2, true_tol=-1.0, max_restarts=0))
    for x in xs:
        x.zero()
    lat.sync()
    t0 = time.perf_counter()
    st = lat.solve_batch(xs, bs, **params)
    lat.sync()
    dt = time.perf_counter() - t0
    return dt, st, xs, bs


PRECISION_WORKLOAD = dict(L=[48, 48, 48, 96], B=[4, 4, 4, 4], eps=0.25, seed=3, kappa=0.137, csw=1.1, tbc=-1,
                          links32=12)


def precision_variant(Q, kappa=None):
    """48^3x96: FP16-stored SAP (FP32 arithmetic; nsap = 4, i.e. 3 sweeps, nm = 6), FP32
    inner (at most 40 iterations), FP64 outer to a 1e-10 FP64 true residual in one solve,
    the outer iteration restarted every 4 iterations (near the critical kappa the
    FP16-stored SAP can make an inner solve diverge; the solve then goes on with the FP32
    SAP and its same 3 sweeps: sap_escalated).  Settings from a sweep at kappa = 0.137:
    maxiter_s 200 / 100 / 50 / 40 / 30 took 204 (nsap 3) / 107 / 76 / 53 / 56 s; the inner
    solves end at 0.1-0.8 of their source at any of these caps.  Its error against an
    FP64-refined continuation: the FP32-SAP mixed solve started from the result and driven
    to a 1e-13 true residual (the preconditioner's precision bounds only the convergence
    rate, not the accuracy of the FP64-refined solution)."""
    w = dict(PRECISION_WORKLOAD, kappa=kappa if kappa is not None else PRECISION_WORKLOAD["kappa"])
    lat, b = setup(Q, w)
    progress("48^3x96 fields ready")
    x = lat.spinor(Q.F64)
    # outer BiCGStab restarted from x every 4 iterations (maxiter 5), as long as the true
    # residual misses 1e-10; a switch to the FP32 SAP holds for the rest of the solve
    prm = dict(SOLVE, sap_precision=Q.F16, nsap=4, nm=6, maxiter=5, max_restarts=60, maxiter_s=40)
    lat.solve(x, b, **dict(prm, maxiter=2, true_tol=-1.0, max_restarts=0))  # warm-up
    x.zero()
    lat.sync()
    t0 = time.perf_counter()
    st = lat.solve(x, b, **prm)
    sec = time.perf_counter() - t0
    progress("FP16-SAP solve: %d outer / %d inner iterations, true residual %.2e, escalated %d, %.0f s"
             % (st["outer_iters"], st["inner_iters"], st["true_residual"], st["sap_escalated"], sec))
    xr = lat.spinor(Q.F64, x.download())
    t1 = time.perf_counter()
    ref = lat.solve(xr, b, **dict(SOLVE, nsap=3, nm=6, tol=1e-14, maxiter=5, max_restarts=60, maxiter_s=40,
                                  true_tol=1e-13))
    ref_sec = time.perf_counter() - t1
    progress("FP64-refined continuation: true residual %.2e, %.0f s" % (ref["true_residual"], ref_sec))
    a, r = x.download(), xr.download()
    err = float(np.linalg.norm(a - r) / np.linalg.norm(r))
    del a, r
    lat.close()
    return {"workload": "48^3x96, eps=0.25 seed 3, kappa=%g, csw=1.1, FP16-stored SAP / FP32 inner / FP64 outer"
                        % w["kappa"],
            "solves_per_s": 1.0 / sec, "seconds": sec, "converged": bool(st["true_residual"] <= 1e-10),
            "true_residual": st["true_residual"], "sap_escalated": st["sap_escalated"],
            "error_vs_fp64_refined": err, "reference_true_residual": ref["true_residual"],
            "iterations": {"outer": st["outer_iters"], "inner": st["inner_iters"], "restarts": st["restarts"]},
            "reference_continuation_seconds": ref_sec}


def eo_operator(Q, lat, reps=20):
    """The even/odd Schur operator mtilde = 1 - kappa^2 T_ee^-1 D_eo T_oo^-1 D_oe
    (qws.cc:551-558, mtilde_vm_ / mtilde_s_) on the even sites of the workload lattice,
    fields taken from the lattice's own: device time per application (host clock over
    `reps` applications between synchronisations) and its HBM rate.  Algorithmic bytes
    per even site: two hop applications, each reading the 8 links of an output site
    (4 of its own, 4 of its neighbours), the clover inverse, the input and writing the
    output; the second also reads in0 -- FP64 (18-real links): 2 x (8 x 144 + 576 + 192 +
    192) + 192 = 4416 B; FP32 with the lattice's 12-real links: 2 x (8 x 48 + 288 + 96 +
    96) + 96 = 1824 B (18-real: 2208 B)."""
    lat.eo_from_lattice()
    vh = lat.V // 2
    nl32 = lat.links32
    out = {"workload": "mtilde on the %d even sites of 32^3x64 (FP64 18-real, FP32 %d-real links; fields from "
                       "the lattice)" % (vh, nl32)}
    rng = np.random.default_rng(3)
    x = rng.standard_normal((vh, 24))
    bps32 = 2 * (8 * nl32 * 4 + 288 + 96 + 96) + 96
    for prec, name, bps in ((Q.F64, "f64", 4416), (Q.F32, "f32", bps32)):
        a = lat.eo_spinor(prec, x)
        o = lat.eo_spinor(prec)
        lat.mtilde(o, a)
        lat.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            lat.mtilde(o, a)
        lat.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        gbs = bps * vh / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "bytes_per_site": bps}
    return out


def eo_solvers(Q, lat, iters=200):
    """The even/odd solvers (bicgstab_vm_, cg_vm_, mcg_vm_ with 4 shifts) on the parity fields
    of the 32^3x64 lattice (FP64, 18-real links), each run for a fixed number of iterations
    (tol 0) from a Gaussian source: ms per iteration, the bandwidth of the mtilde applications
    that dominate it (two mtilde-sized operator passes per iteration for all three: M p and
    M s; mtilde and its adjoint; 4416 B per even site each, as in eo_operator) and the
    relative residual reached, recomputed with mtilde on the downloaded solution."""
    vh = lat.V // 2
    rng = np.random.default_rng(5)
    bh = rng.standard_normal((vh, 24))
    b = lat.eo_spinor(Q.F64, bh)
    out = {"workload": "even/odd solvers on the %d even sites of 32^3x64, FP64, %d iterations each from a Gaussian "
                       "source" % (vh, iters)}
    op_bytes = 2 * 4416 * vh

    def resid(x, which):
        r = lat.eo_spinor(Q.F64)
        lat.mtilde(r, x, which=which)
        d = r.download()
        return float(np.linalg.norm(bh - d) / np.linalg.norm(bh))

    def timed(f):
        lat.sync()
        t0 = time.perf_counter()
        it = f()
        lat.sync()
        return time.perf_counter() - t0, it

    x = lat.eo_spinor(Q.F64)
    sec, it = timed(lambda: lat.eo_bicgstab(x, b, maxiter=iters, tol=0.0))
    out["bicgstab_vm"] = {"ms_per_iteration": sec * 1e3 / it, "iterations": it,
                          "operator_GBs": op_bytes * it / sec / 1e9, "relative_residual": resid(x, 0)}
    x = lat.eo_spinor(Q.F64)
    sec, it = timed(lambda: lat.eo_cg(x, b, maxiter=iters, tol=0.0))
    out["cg_vm"] = {"ms_per_iteration": sec * 1e3 / it, "iterations": it,
                    "operator_GBs": op_bytes * it / sec / 1e9, "relative_residual": resid(x, 2)}
    shifts = [0.0, 0.01, 0.1, 1.0]
    res = {}
    sec, (xs, it) = timed(lambda: lat.eo_mcg(b, shifts, maxiter=iters, tol=0.0))
    out["mcg_vm"] = {"shifts": shifts, "ms_per_iteration": sec * 1e3 / it, "iterations": it,
                     "operator_GBs": op_bytes * it / sec / 1e9, "relative_residual_shift0": resid(xs[0], 2)}
    return out


def qws_host_layout(a, kind, L, vlen):
    """Plain fields (site x + Lx (y + Ly (z + Lz t))) -> QWS's host AoSoA (qws.h:111-192,
    main.cc:426-429,485-488): x split into two domains of Lx/2, vector i = xv + nxv (y + Ly
    (z + Lz t)), lane = x % vlen; spinors [i][col][spin][ri][lane], links
    [i + vol mu + 4 vol d][row][col][ri][lane], clover [i + vol d][chirality][k][lane]."""
    nx, ny, nz, nt = L
    nxv = nx // 2 // vlen
    if kind == "spinor":
        p = np.asarray(a).reshape(nt, nz, ny, 2, nxv, vlen, 3, 4, 2)
        return np.ascontiguousarray(p.transpose(3, 0, 1, 2, 4, 6, 7, 8, 5))
    if kind == "links":
        p = np.asarray(a).reshape(4, nt, nz, ny, 2, nxv, vlen, 3, 3, 2)
        return np.ascontiguousarray(p.transpose(4, 0, 1, 2, 3, 5, 7, 8, 9, 6))
    p = np.asarray(a).reshape(nt, nz, ny, 2, nxv, vlen, 2, 36)
    return np.ascontiguousarray(p.transpose(3, 0, 1, 2, 4, 6, 7, 5))


def dropin_solve(Q, reps=2):
    """The workload through QWS's own entry point bicgstab_dd_mix_ (C ABI, host AoSoA
    arrays in and out): fields written into glud/glus/clvd/clvs as a QWS driver does, 4^4
    SAP blocks (qws_mi355x_set_sap_blocks_), the solve's settings without the restarts on
    the true residual.  The first call also uploads the fields; the timed calls include
    the host<->device copies and layout conversions of x and b (the PCIe-inclusive rate)."""
    import ctypes
    w = WORKLOAD
    lat, b = setup(Q)
    U, C, bp = lat.gauge(), lat.clover_field(), b.download()
    lat.close()
    del b
    L = Q.lib()
    nx, ny, nz, nt = w["L"]
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cd = lambda v: ctypes.byref(ctypes.c_double(v))
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, w["tbc"]), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(ci(nx), ci(ny), ci(nz), ci(nt), keep[0], keep[1], ctypes.byref(keep[2]), ctypes.byref(keep[3]),
                keep[4])

    def fill(name, arr):
        dst = np.ctypeslib.as_array((ctypes.c_byte * arr.nbytes).from_address(ctypes.c_void_p.in_dll(L, name).value))
        dst[:] = arr.view(np.int8).ravel()

    fill("glud", qws_host_layout(U, "links", w["L"], 8))
    fill("glus", qws_host_layout(U.astype(np.float32), "links", w["L"], 16))
    del U
    fill("clvd", qws_host_layout(C, "clover", w["L"], 8))
    fill("clvs", qws_host_layout(C.astype(np.float32), "clover", w["L"], 16))
    del C
    for name, v in (("kappa", w["kappa"]), ("mkappa", -w["kappa"]), ("kappa2", w["kappa"] ** 2)):
        ctypes.c_double.in_dll(L, name).value = v
    L.qws_mi355x_fields_modified_()
    L.qws_mi355x_set_sap_blocks_((ctypes.c_int * 4)(*w["B"]))
    if w["links32"] == 12:  # SU(3) links: 12-real FP32 storage on the GPU, as the workload
        L.qws_mi355x_set_links12_(ci(1))
    bh = qws_host_layout(bp, "spinor", w["L"], 8)
    x = np.zeros_like(bh)
    it = ctypes.c_int(0)
    args = lambda: (x.ctypes.data_as(ctypes.c_void_p), bh.ctypes.data_as(ctypes.c_void_p), cd(SOLVE["tol"]),
                    ctypes.byref(it), ci(SOLVE["maxiter"]), cd(SOLVE["tol_s"]), ci(SOLVE["maxiter_s"]),
                    ci(SOLVE["nsap"]), ci(SOLVE["nm"]))
    t0 = time.perf_counter()
    L.bicgstab_dd_mix_(*args())
    first = time.perf_counter() - t0
    times = []
    for _ in range(reps):
        x[:] = 0
        t0 = time.perf_counter()
        L.bicgstab_dd_mix_(*args())
        times.append(time.perf_counter() - t0)
    r = np.zeros_like(bh)
    L.ddd_d_(r.ctypes.data_as(ctypes.c_void_p), x.ctypes.data_as(ctypes.c_void_p))
    res = float(np.sqrt(np.sum((bh - r) ** 2) / np.sum(bh ** 2)))
    L.qws_finalize_()
    t = float(np.mean(times))
    return {"workload": "bicgstab_dd_mix_ through the QWS C ABI on the same 32^3x64 fields and source, host "
                        "AoSoA x and b, 4^4 SAP blocks, %d-real FP32 links" % w["links32"],
            "solves_per_s": 1.0 / t, "s_per_solve": t,
            "first_call_s": first, "outer_iters": it.value, "true_residual": res}


SERVE_VOLUMES = ([8, 8, 8, 8], [16, 16, 16, 32], [24, 24, 24, 48], [32, 32, 32, 64])


def uneven_batch(Q, params, fused_only=0):
    """16^3x32, 24^3x48 and 32^3x64 (links exp(0.3 iH), seeds 9-11; Gaussian sources with
    seeds 4-6; kappa drawn from U[0.130, 0.137]) solved at once with every SAP half-sweep
    fused over the three lattices (qwsg_solve_multi), against the same three solves one
    after another."""
    rng = np.random.default_rng(4)
    kap = [float(rng.uniform(0.130, 0.137)) for _ in range(3)]
    lats, parts = [], []
    for i, L in enumerate(SERVE_VOLUMES[1:]):
        lat = Q.Lattice(L, WORKLOAD["B"], WORKLOAD["tbc"], WORKLOAD["links32"], kappa=kap[i])
        lat.generate_gauge(WORKLOAD["eps"], 9 + i)
        lat.compute_clover(WORKLOAD["csw"])
        lats.append(lat)
        parts.append((lat, [lat.spinor(Q.F64)], [lat.gaussian(Q.F64, 4 + i)]))

    def fused():
        for _, xs, _ in parts:
            xs[0].zero()
        for lat in lats:
            lat.sync()
        t0 = time.perf_counter()
        st = Q.solve_multi(parts, **params)
        return time.perf_counter() - t0, st

    def separate():
        t0 = time.perf_counter()
        st = []
        for lat, xs, bs in parts:
            xs[0].zero()
            st.append([lat.solve(xs[0], bs[0], **params)])
            lat.sync()
        return time.perf_counter() - t0, st

    if fused_only:
        # only the fused solves (for kernel traces): warm-up, then `fused_only` timed solves
        fused()
        ts = [fused()[0] for _ in range(fused_only)]
        st = fused()[1]
        for lat in lats:
            lat.close()
        return {"fused_seconds": ts, "kappa": kap, "inner_iters": [s[0]["inner_iters"] for s in st],
                "outer_iters": [s[0]["outer_iters"] for s in st]}
    fused()  # warm-up (buffers, descriptor tables, graphs)
    separate()
    t_sep, st_sep = separate()
    t_fused, st = fused()
    out = {"workload": "16^3x32, 24^3x48, 32^3x64 (one Gaussian source each, seeds 4-6, kappa "
                       + ", ".join("%.4f" % k for k in kap) + "), SAP half-sweeps fused over the three lattices",
           "solves_per_s": 3 / t_fused, "seconds": t_fused, "separate_seconds": t_sep,
           "fused_vs_separate": t_sep / t_fused,
           "c2_equiv_solves_per_s": sum(int(np.prod(L)) for L in SERVE_VOLUMES[1:]) / int(np.prod(WORKLOAD["L"]))
           / t_fused,
           "true_residual_max": max(s[0]["true_residual"] for s in st),
           "same_iterations_as_separate": all(a[0]["inner_iters"] == b[0]["inner_iters"] for a, b in zip(st, st_sep))}
    # the fused solutions against all-FP64 solves of the same sources (outside the timing)
    out["error_vs_fp64"] = {"x".join(map(str, lat.L)): error_vs_fp64(Q, lat, xs[0], bs[0])["error"]
                            for lat, xs, bs in parts}
    for lat in lats:
        lat.close()
    return out


def serve_trace(rate, n, seed=7):
    """Poisson arrivals at `rate` req/s; volume uniform over SERVE_VOLUMES (field i has
    volume i), kappa ~ U[0.125, 0.137], 1..12 Gaussian right-hand sides."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    return [dict(field=int(rng.integers(4)), kappa=float(rng.uniform(0.125, 0.137)), nrhs=int(rng.integers(1, 13)),
                 seed=1000 + i, arrival=float(t[i])) for i in range(n)]


def serve_stats(res, trace):
    lat_ms = np.array([(r["finish"] - r["arrival"]) * 1e3 for r in res])
    svc_ms = np.array([(r["finish"] - r["admit"]) * 1e3 for r in res])
    span = max(r["finish"] for r in res) - min(r["arrival"] for r in res)
    rhs = sum(t["nrhs"] for t in trace)
    # right-hand sides weighted by volume, in 32^3x64 solves
    c2 = sum(t["nrhs"] * int(np.prod(SERVE_VOLUMES[t["field"]])) for t in trace) / int(np.prod(WORKLOAD["L"]))
    return {"requests": len(res), "rhs": rhs, "solves_per_s": rhs / span, "requests_per_s": len(res) / span,
            "c2_equiv_solves_per_s": c2 / span,
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
            "service_p50_ms": float(np.percentile(svc_ms, 50)), "service_p99_ms": float(np.percentile(svc_ms, 99)),
            "true_residual_max": max(r["true_residual_max"] for r in res)}


def make_server(Q, slots=4):
    fields = [(L, [4, 4, 4, 4], 0.3, 8 + i) for i, L in enumerate(SERVE_VOLUMES)]
    return Q.Server(fields, links32=WORKLOAD["links32"], tbc=WORKLOAD["tbc"], csw=WORKLOAD["csw"], slots=slots,
                    **SOLVE)


CPU_REFERENCE_SOLVE = os.path.join(ROOT, "profiles", "cpu_reference_32c64.json")


def cpu_baseline(Q, lat, stats, workdir):
    """QWS reference (oracle/_ref/libqws_ref.so) on this host's cores.

    QWS's own full solve of this workload (bicgstab_dd_mix_ to the outer tolerance in its
    single-rank geometry, two SAP domains of 16 x 32 x 32 x 64) was measured once by
    oracle/cpu_reference_solve.py on another host (profiles/cpu_reference_32c64.json:
    its wall time, outer iterations, true residual and its time per inner iteration).
    Here a bounded sample repeats that inner-iteration timing (bicgstab_precdd_s_, 2
    fixed iterations, same nsap/nm and fields) on this host's cores, and the full solve's
    time is scaled by the ratio of the two per-iteration times.  Without the fixture the
    sample is scaled by the GPU solve's inner iteration count instead (a model)."""
    from oracle import ref_worker as RW

    if not RW.available():
        return None
    U = lat.gauge()
    C = lat.clover_field()
    b = lat.gaussian(Q.F64, WORKLOAD["seed"]).download()
    nloc = WORKLOAD["L"]
    fin = os.path.join(workdir, "in.npz")
    fout = os.path.join(workdir, "out.npz")
    sample_iters = 2
    np.savez(fin, nloc=np.array(nloc), kappa=np.array(WORKLOAD["kappa"]), U=U, C=C, b=b, x0=np.zeros_like(b),
             ops=np.array(["time_inner"]), nsap=np.array(SOLVE["nsap"]), nm=np.array(SOLVE["nm"]), tol=np.array(-1.0),
             maxiter=np.array(2), tol_s=np.array(-1.0), maxiter_s=np.array(sample_iters))
    del U, C
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.perf_counter()
    p = subprocess.run([sys.executable, "-m", "oracle.ref_worker", fin, fout], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=900)
    if p.returncode != 0:
        print("cpu baseline failed:", p.stderr[-2000:], file=sys.stderr)
        return None
    res = dict(np.load(fout))
    t_iter = float(res["time_inner"]) / sample_iters
    wall = time.perf_counter() - t0
    if os.path.exists(CPU_REFERENCE_SOLVE):
        ref = json.load(open(CPU_REFERENCE_SOLVE))
        t_solve = ref["solve_seconds"] * t_iter / ref["inner_iteration_seconds"]
        sample = (f"{sample_iters} QWS inner iterations (bicgstab_precdd_s_, nsap={SOLVE['nsap']}, nm={SOLVE['nm']}) "
                  f"on the same 32^3x64 fields here: {t_iter:.3f} s/iteration on {cores} threads; QWS's full solve "
                  f"of this workload measured on {ref['threads']} threads of a {ref['cpu']}: {ref['solve_seconds']:.1f} s, "
                  f"{ref['outer_iters']} outer iterations, true residual {ref['true_residual']:.1e}, "
                  f"{ref['inner_iteration_seconds']:.3f} s/inner iteration there; scaled by the per-iteration ratio "
                  f"(QWS single-rank SAP domains 16x32x32x64); sample wall {wall:.0f} s")
    else:
        inner = np.mean([s["inner_iters"] for s in stats])
        t_solve = t_iter * inner
        sample = (f"model: {sample_iters} QWS inner iterations here ({t_iter:.3f} s/iteration) x the GPU solve's "
                  f"{inner:.0f} inner iterations; sample wall {wall:.0f} s")
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": cores, "kind": "reference", "sample": sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--dslash-sweep", type=int, default=0,
                    help="time the full operator for every block order and kernel variant (N launches each)")
    ap.add_argument("--uneven-only", type=int, default=0,
                    help="only N fused solves of the uneven batch (16^3x32, 24^3x48, 32^3x64), for kernel traces")
    ap.add_argument("--mfma-trial", type=int, default=0,
                    help="time the multi-RHS clover product on the VALU and on MFMA (N launches each)")
    ap.add_argument("--dslash-only", type=int, default=0, help="only time N dslash launches (for counter runs)")
    ap.add_argument("--dropin", action="store_true",
                    help="the workload through bicgstab_dd_mix_ with host arrays (PCIe-inclusive rate)")
    ap.add_argument("--eo-only", type=int, default=0,
                    help="only the even/odd measurements: N applications of mtilde (FP64, FP32), the solvers")
    ap.add_argument("--sap-only", type=int, default=0,
                    help="only time N preconditioner applications of fused SAP launches (for counter runs)")
    ap.add_argument("--sweep", action="store_true", help="print solve statistics for a grid of solver settings")
    ap.add_argument("--serve", action="store_true",
                    help="serving sweep: Poisson rates 1..200 req/s over four resident fields (p50/p99 latency)")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-slots", type=int, default=4, help="concurrent requests of the server (1..16)")
    ap.add_argument("--serve-rates", default="1,2,5,10,20,50,100,200", help="offered rates of the sweep (req/s)")
    ap.add_argument("--only-plain", action="store_true", help="only the plain solves and the D roofline (profiling)")
    ap.add_argument("--blocks", action="store_true",
                    help="time one solve of the workload per SAP block shape (qwsg_autotune_blocks)")
    ap.add_argument("--precision", action="store_true",
                    help="the 48^3x96 FP16-SAP precision variant and its error against an FP64-refined solution")
    ap.add_argument("--precision-kappa", type=float, default=None, help="kappa of the precision variant")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="only run N device-to-device copies of 1 GiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this solver runs on one GPU; sharding over several GPUs is not built",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.copy_calib:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters give the
        # factor between counter units and bytes
        import torch

        n = 1 << 28  # 2^28 floats = 1 GiB
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(a.copy_calib):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each": 4 * n, "copies": a.copy_calib}))
        return
    import qws_amd as Q

    if a.precision:
        print(json.dumps(precision_variant(Q, a.precision_kappa)), flush=True)
        return
    if a.serve:
        srv = make_server(Q, slots=a.serve_slots)
        # warm-up: the same requests (the trace's shapes do not depend on the rate) at once,
        # which fills the slot pool
        srv.run(serve_trace(1000.0, a.serve_requests))
        # 1 and 2 req/s lie below the server's capacity (~4.2 requests/s of this mix with
        # 4 slots); 5..200 is the configured sweep, at or above it
        for rate in [float(v) for v in a.serve_rates.split(",")]:
            tr = serve_trace(rate, a.serve_requests)
            res = srv.run(tr)
            print(json.dumps({"rate_req_per_s": rate, "slots": a.serve_slots, **serve_stats(res, tr)}), flush=True)
        srv.close()
        return
    if a.uneven_only:
        print(json.dumps(uneven_batch(Q, SOLVE, fused_only=a.uneven_only)), flush=True)
        return
    if a.mfma_trial:
        # the multi-RHS clover product, VALU vs MFMA, on 2^20 sites x 16 RHS
        for nrep in (1, 4, 7, 16, 64):
            ms_v, ms_m, diff = Q.time_clover_rhs(1 << 20, nrep, a.mfma_trial)
            useful = (1 << 20) * 16 * nrep * 2 * 144 * 2  # real-form 12x12 products, flops
            print(json.dumps({"nrep": nrep, "valu_ms": ms_v, "mfma_ms": ms_m, "mfma_speedup": ms_v / ms_m,
                              "valu_TFLOPs": useful / (ms_v * 1e-3) / 1e12,
                              "mfma_TFLOPs": useful / (ms_m * 1e-3) / 1e12, "max_rel_diff": diff}), flush=True)
        return
    if a.dropin:
        print(json.dumps(dropin_solve(Q)))
        return
    lat, b = setup(Q)
    V = lat.V
    if a.blocks:
        # 256-site (two workgroups per CU) and 512-site blocks; one untimed solve each first
        shapes = [[4, 4, 4, 4], [8, 4, 4, 2], [8, 4, 2, 4], [8, 4, 4, 4], [4, 4, 4, 8]]
        best, sec = lat.autotune_blocks(b, shapes, **SOLVE)
        print(json.dumps({"block_shapes": shapes, "seconds_per_solve": sec, "best": best}), flush=True)
        return
    if a.dslash_sweep:
        # block order x kernel variant of the full operator (FP32), and FP64 D per order
        for order in (0, 1):
            lat.set_dslash_order(order)
            for k in (0, 3, 4):
                lat.set_dslash_kernel(k)
                ms = lat.time_dslash(Q.F32, a.dslash_sweep)
                print(json.dumps({"order": order, "kernel": k, "f32_ms": ms,
                                  "GBs": V * BYTES_PER_SITE_F32_12 / (ms * 1e-3) / 1e9}), flush=True)
            ms = lat.time_dslash(Q.F64, a.dslash_sweep)
            print(json.dumps({"order": order, "f64_ms": ms, "GBs": V * 1536 / (ms * 1e-3) / 1e9}), flush=True)
        # FP64 kernel variants (QWS_MI355X_DSLASH_D, read when a lattice is created):
        # 0 global gathers, 2 / 3 LDS-staged for 2 / 3 waves per SIMD; index order
        lat.close()
        for kd in ("0", "2", "3"):
            os.environ["QWS_MI355X_DSLASH_D"] = kd
            lk, _ = setup(Q)
            ms = lk.time_dslash(Q.F64, a.dslash_sweep)
            print(json.dumps({"order": 0, "f64_kernel": int(kd), "f64_ms": ms,
                              "GBs": V * 1536 / (ms * 1e-3) / 1e9}), flush=True)
            lk.close()
        os.environ.pop("QWS_MI355X_DSLASH_D")
        return
    if a.dslash_only:
        ms = lat.time_dslash(Q.F32, a.dslash_only)
        print(json.dumps({"dslash_ms": ms, "GBs": V * BYTES_PER_SITE_F32_12 / (ms * 1e-3) / 1e9}))
        return
    if a.eo_only:
        print(json.dumps({"eo": eo_operator(Q, lat, a.eo_only), "eo_solvers": eo_solvers(Q, lat)}), flush=True)
        return
    if a.sap_only:
        ms = lat.time_sap(Q.F32, a.sap_only, SOLVE["nsap"], SOLVE["nm"])
        fl = sap_flops_per_site(WORKLOAD["B"], SOLVE["nsap"], SOLVE["nm"]) * V / 2
        print(json.dumps({"sap_ms_per_launch": ms, "TFLOPs": fl / (ms * 1e-3) / 1e12}))
        return
    if a.sweep:
        # the neighbourhood of the chosen settings; two solves each (seconds per solve)
        t_start = time.perf_counter()
        run_solves(Q, lat, b, 1, SOLVE)  # warm-up
        for nsap in (1, 2, 3):
            for nm in (4, 5, 6, 7, 8):
                for tol_s in (1e-3, 3e-3):
                    if time.perf_counter() - t_start > 240:
                        print(json.dumps({"sweep": "stopped at its time budget"}), flush=True)
                        return
                    prm = dict(SOLVE, nsap=nsap, nm=nm, tol_s=tol_s)
                    dt, st, _ = run_solves(Q, lat, b, 2, prm)
                    print(json.dumps({"nsap": nsap, "nm": nm, "tol_s": tol_s, "s": round(dt / 2, 4), **st[0]}),
                          flush=True)
        return

    # warm-up and timed solves
    run_solves(Q, lat, b, a.warmup, SOLVE)
    dt, stats, x = run_solves(Q, lat, b, a.steps, SOLVE)
    ms_per = dt / a.steps * 1e3
    worst = max(s["true_residual"] for s in stats)
    # rooflines: HIP events on the library stream, same fields
    d_kernel = lat.autotune_dslash(20)
    ms_d = lat.time_dslash(Q.F32, 50)
    achieved = V * BYTES_PER_SITE_F32_12 / (ms_d * 1e-3) / 1e9
    ms_s = lat.time_sap(Q.F32, 10, SOLVE["nsap"], SOLVE["nm"])
    flops_s = sap_flops_per_site(WORKLOAD["B"], SOLVE["nsap"], SOLVE["nm"]) * V / 2
    achieved_s = flops_s / (ms_s * 1e-3) / 1e12
    # practical HBM ceiling: streaming kernels with D's read:write mix (6:1) over six arrays
    # and over one array read region by region, and a copy; D is compared with the best
    stream = {"read6_write1_GBs": Q.stream_bandwidth(256 << 20, 6, 20),
              "read6_write1_one_stream_GBs": Q.stream_bandwidth(256 << 20, 6, 20, one_stream=True),
              "copy_GBs": Q.stream_bandwidth(1 << 30, 1, 10)}
    stream_best = max(stream.values())

    def traffic(name):
        tfile = os.path.join(ROOT, "profiles", name)
        return json.load(open(tfile)).get("bytes_per_launch") if os.path.exists(tfile) else None
    line = {
        "metric": "solves/s at 1e-10 residual (batched/served vs plain); dslash GB/s; p99 ms",
        "value": 1e3 / ms_per, "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64/f32",
        "data": "synthetic (Philox SU(3) links exp(0.3iH), clover from links, Gaussian source)",
        "config": {"workload": "plain single-RHS solve, 32^3x64, 4^4 SAP blocks, kappa=0.1365, csw=1.0, "
                               "FP64 outer / FP32 inner to 1e-10 FP64 true residual",
                   "lattice": WORKLOAD["L"], "sap_block": WORKLOAD["B"], **SOLVE},
        "true_residual_max": worst,
        "iterations": {"outer": stats[-1]["outer_iters"], "inner": stats[-1]["inner_iters"],
                       "restarts": stats[-1]["restarts"]},
        "dslash_GBs": achieved,
        # compulsory bytes of the inner solve's launches per wall second (the FP64 outer
        # iterations' traffic is not counted: a lower bound)
        "effective_GBs": stats[-1]["inner_iters"] * V * inner_iteration_bytes_per_site(SOLVE["nsap"]) / (
            ms_per * 1e-3) / 1e9,
        "roofline": {"bound": "valu", "achieved": achieved_s, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_s / FP32_PEAK_TFLOPS, "traffic": traffic("sap_f32_32c64_traffic.json"),
                     "kernel": "k_sap_block<float,12> (fused SAP half-sweep, FP32, %.0f flop/site algorithmic, "
                               "V/2 sites per launch)" % (flops_s / (V / 2)),
                     "ms_per_launch": ms_s,
                     "note": "FP32 vector arithmetic (3x3 complex link products: no MFMA shape); the peak "
                             "counts packed FP32 FMA, which this build does not emit (unpacked peak 78.6)"},
        "roofline_d": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic("dslash_f32_32c64_traffic.json"),
                       "kernel": ("k_dslash<float,12>" if d_kernel == 0 else "k_dslash_lds<12,%d>" % d_kernel)
                                 + " (FP32 D, 672 B/site algorithmic; kernel chosen by autotune_dslash)",
                       "ms_per_launch": ms_d, "stream_ceiling": stream,
                       "frac_of_stream": achieved / stream_best},
        "cpu_baseline": None,
    }
    if a.only_plain:
        # (profiling: the solves and the operator timings only)
        print(json.dumps(line), flush=True)
        return
    line["error_vs_fp64"] = error_vs_fp64(Q, lat, x, b)
    progress("plain solves, dslash and the all-FP64 comparison done")
    line["eo"] = eo_operator(Q, lat)
    line["eo_solvers"] = eo_solvers(Q, lat)
    progress("even/odd operator and solvers done")
    # batched: the 12 point sources on the same field, one batch
    dt_b, st_b, xs_b, bs_b = batched_point_sources(Q, lat, SOLVE)
    line["batched"] = {"workload": "12 spin-colour point sources at the origin, one batched solve, 32^3x64",
                       "solves_per_s": 12 / dt_b, "vs_plain": (12 / dt_b) / line["value"],
                       "true_residual_max": max(s["true_residual"] for s in st_b),
                       "inner_iters": [s["inner_iters"] for s in st_b],
                       "error_vs_fp64": {"source %d" % k: error_vs_fp64(Q, lat, xs_b[k], bs_b[k]) for k in (0, 7)}}
    del xs_b, bs_b
    progress("batched point sources done")
    # uneven batch: three volumes in one fused solve
    line["uneven"] = uneven_batch(Q, SOLVE)
    progress("uneven batch done")
    # serving: four resident fields of different volume, continuous batching
    srv = make_server(Q, slots=4)
    srv.run(serve_trace(1000.0, 16))  # warm-up: the same requests at once (fills the slot pool)
    tr = serve_trace(20.0, 16)
    res = srv.run(tr)
    line["served"] = {"workload": "Poisson 20 req/s, 16 requests, volumes 8^4/16^3x32/24^3x48/32^3x64, "
                                  "kappa U[0.125,0.137], 1-12 RHS, up to 4 requests in the running batch",
                      **serve_stats(res, tr)}
    line["served"]["c2_equiv_vs_plain"] = line["served"]["c2_equiv_solves_per_s"] / line["value"]
    srv.close()
    progress("served trace done")
    # the precision variant (48^3x96, eps 0.25, csw 1.1, FP16-stored SAP) at its kappa = 0.137
    # (near critical: about a minute, and the FP64-refined continuation) and at kappa = 0.13
    line["precision"] = precision_variant(Q)
    progress("precision variant (kappa 0.137) done")
    line["precision_kappa_0.13"] = precision_variant(Q, 0.13)
    progress("precision variant (kappa 0.13) done")
    if not a.no_cpu_baseline:
        progress("timing the CPU reference")
        with tempfile.TemporaryDirectory() as td:
            line["cpu_baseline"] = cpu_baseline(Q, lat, stats, td)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
