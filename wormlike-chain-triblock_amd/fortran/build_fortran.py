"""Build the Fortran binding (module wormlike_MDE), the SCFT field update (module
scft_iterate), the field files (module field_io) and the SCFT driver program with ROCm
flang, linked against libwlc_mde.so (rpath to the package directory); field_io_check,
a GPU-free round trip of the field files; and fe_check, the GPU-free field energy."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
BUILD = os.path.join(PKG, "_build", "fortran")
EXE = os.path.join(HERE, "scft_driver")
CHECK = os.path.join(HERE, "field_io_check")
FE_CHECK = os.path.join(HERE, "fe_check")


def find_flang():
    for c in ("amdflang", "flang", "/opt/rocm/bin/amdflang", "/opt/rocm/lib/llvm/bin/flang"):
        p = shutil.which(c) or (c if os.path.exists(c) else None)
        if p:
            return p
    return None


def build():
    fc = find_flang()
    if fc is None:
        print("build_fortran: no Fortran compiler found; skipping the Fortran binding")
        return None
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    steps = []
    for src in ("scft_energy.f90", "wormlike_mde_gpu.f90", "scft_iterate.f90", "field_io.f90",
                "scft_driver.f90", "field_io_check.f90", "fe_check.f90"):
        obj = os.path.join(BUILD, src.replace(".f90", ".o"))
        steps.append([fc, "-O2", "-module-dir", BUILD, "-I", BUILD, "-c", os.path.join(HERE, src),
                      "-o", obj])
        objs.append(obj)
    tmp, tmp_check, tmp_fe = EXE + ".tmp", CHECK + ".tmp", FE_CHECK + ".tmp"
    steps.append([fc, "-o", tmp] + objs[:5] + [f"-L{PKG}", "-lwlc_mde", "-Wl,-rpath,$ORIGIN/.."])
    steps.append([fc, "-o", tmp_check, objs[3], objs[5]])
    steps.append([fc, "-o", tmp_fe, objs[0], objs[6]])
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=BUILD)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed ({' '.join(cmd)}):\n{r.stderr[-4000:]}")
    os.replace(tmp, EXE)
    os.replace(tmp_check, CHECK)
    os.replace(tmp_fe, FE_CHECK)
    print(f"built {EXE}, {CHECK}, {FE_CHECK}")
    return EXE


if __name__ == "__main__":
    build()
    sys.exit(0)
