"""Build libwlc_mde.so (HIP, gfx950) in-tree.

Every translation unit is compiled by hipcc in parallel into _build/, then linked
into libwlc_mde.so beside this file.  wlc_zsolve.hip is compiled once per Lmax
(0..8), each object carrying that L's kernel-B instantiation.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwlc_mde.so")
ARCH = os.environ.get("WLC_OFFLOAD_ARCH", "gfx950")
MAX_L = 8

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep hipcc from emitting packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _units():
    units = [
        ("wlc_mde.hip", "wlc_mde.o", []),
        ("wlc_batch.hip", "wlc_batch.o", []),
        ("wlc_unet.hip", "wlc_unet.o", []),
        ("wlc_plane.hip", "wlc_plane.o", []),
        ("wlc_dispatch.cpp", "wlc_dispatch.o", ["-x", "hip"]),
        ("sh_setup.cpp", "sh_setup.o", []),
        ("wlc_tune.cpp", "wlc_tune.o", []),
    ]
    for L in range(MAX_L + 1):
        units.append(("wlc_zsolve.hip", f"wlc_zsolve_L{L}.o", [f"-DWLC_L={L}"]))
    return units


def _stale(unit):
    """True when the object is missing or older than its source or any header of csrc/ and
    include/ (incremental development builds; build() without incremental recompiles all)."""
    src, obj, _ = unit
    o = os.path.join(BUILD, obj)
    if not os.path.exists(o):
        return True
    t = os.path.getmtime(o)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC)
                                        if f.endswith((".h", ".inc"))]
    inc = os.path.join(os.path.dirname(HERE), "include")
    deps += [os.path.join(inc, f) for f in os.listdir(inc)]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(unit):
    src, obj, extra = unit
    cmd = [HIPCC] + FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src} ({' '.join(extra)}):\n{r.stderr[-4000:]}")
    return obj


def gen_tables():
    """Regenerate csrc/wlc_tables.inc from the host set-up code (g++)."""
    exe = os.path.join(BUILD, "gen_tables")
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-o", exe, os.path.join(CSRC, "gen_tables.cpp"),
           os.path.join(CSRC, "sh_setup.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gen_tables build failed:\n{r.stderr[-4000:]}")
    out = os.path.join(BUILD, "wlc_tables.inc")
    r = subprocess.run([exe, out], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gen_tables failed:\n{r.stderr}")
    dst = os.path.join(CSRC, "wlc_tables.inc")
    new = open(out).read()
    if not os.path.exists(dst) or open(dst).read() != new:
        with open(dst, "w") as f:
            f.write(new)


def build(jobs=None, verbose=True, incremental=False):
    os.makedirs(BUILD, exist_ok=True)
    gen_tables()
    units = _units()
    todo = [u for u in units if _stale(u)] if incremental else units
    jobs = jobs or min(max(1, len(todo)), max(1, min(os.cpu_count() or 1, 16)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [u[1] for u in units]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + \
        [os.path.join(BUILD, o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
