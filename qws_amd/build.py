"""Build libqws_mi355x.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m qws_amd.build [--jobs N]

Every source in qws_amd/csrc is compiled as HIP for gfx950 and linked into
qws_amd/lib/libqws_mi355x.so.  Packed-FP32 instructions are disabled for the
device code (-target-feature -packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libqws_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QWS_OFFLOAD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-Wno-unknown-pragmas", "-Wno-unused-result"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, p.returncode, p.stdout + p.stderr


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = sources()
    objs = []
    failed = False
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, out in ex.map(compile_one, srcs):
            objs.append(obj)
            msgs = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if rc != 0:
                failed = True
                print(f"[build] FAILED {os.path.basename(src)}\n" + "\n".join(msgs), file=sys.stderr)
            elif verbose and msgs:
                print(f"[build] {os.path.basename(src)}:\n" + "\n".join(msgs), file=sys.stderr)
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
