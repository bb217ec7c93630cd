"""Build recipe of the native library: hipcc for gfx950, in-tree output.

    python -m sharpclaw_amd.build        -> sharpclaw_amd/libsharpclaw_amd.so

The library is the C ABI of include/sharpclaw_amd.h.  It is compiled
ahead of time here (hipcc cross-compiles without a GPU) and loaded from the
package directory at run time.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsharpclaw_amd.so")
SOURCES = ["sc_kernels.hip", "sc_batch.hip", "sc_batch2d.hip", "sc_capi.cpp", "sc_serve.cpp",
           "sc_noalloc.hip"]
ARCH = os.environ.get("SC_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        # packed-FP32 VALU forms are kept out of the generated code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unknown-pragmas",
        f"-I{os.path.join(ROOT, 'include')}",
    ]


def build(verbose=False):
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(bdir, src + ".o")
        cmds.append([hipcc(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + flags())
        objs.append(obj)
    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for f in [pool.submit(subprocess.check_call, cmd) for cmd in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
# sample applications: (ndim, their hook sources under fortran/apps)
FAPPS = {
    "adv1d": (1, ["setprob_advection.f90", "qinit1.f90"]),
    "ac1d": (1, ["setprob_acoustics.f90", "qinit1.f90"]),
    "eu1d": (1, ["setprob_euler.f90", "qinit1.f90"]),
    "sw1d": (1, ["setprob_sw.f90", "qinit1.f90"]),
    "eu2d": (2, ["setprob_euler.f90", "qinit2.f90"]),
    "sw2d": (2, ["setprob_sw.f90", "qinit2.f90"]),
    # user bc (mthbc 0), src and b4step linked in place of the defaults
    "ac1du": (1, ["setprob_acoustics.f90", "qinit1.f90", "userhooks_ac1d.f90"]),
}
# default hooks, one routine per archive member, so that an application's
# own routine replaces exactly its default (link-time substitution, as the
# reference's application Makefiles omit the library's bc1.f90 or src1.f90)
FDEFAULTS = {
    1: ["b4step.f90", "setaux.f90", "skip_data_header.f90", "bc1.f90", "src1.f90"],
    2: ["b4step.f90", "setaux.f90", "skip_data_header.f90", "bc2.f90", "src2.f90"],
}


def flang():
    for c in (os.environ.get("FC"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose=False):
    """The Fortran host driver (fortran/xsclaw_gpu.f90 + the ISO_C_BINDING
    module) linked with each sample application's hooks into
    fortran/_build/xsclaw_gpu_<app>, with an rpath to this package."""
    fc = flang()
    if fc is None:
        if verbose:
            print("flang not found: Fortran driver not built", flush=True)
        return []
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd, cwd=bdir)

    run([fc, "-O2", "-c", os.path.join(FORTRAN, "clawparams.f90"), "-o", "clawparams.o"])
    run([fc, "-O2", "-c", os.path.join(FORTRAN, "sharpclaw_gpu.f90"), "-o", "sharpclaw_gpu.o"])
    drv = ["clawparams.o", "sharpclaw_gpu.o"]
    for src in ("xsclaw_gpu.f90", "host_bc1.f90", "host_bc2.f90"):
        obj = src[:-4] + ".o"
        run([fc, "-O2", "-c", os.path.join(FORTRAN, src), "-o", obj])
        drv.append(obj)
    for nd, srcs in FDEFAULTS.items():
        objs = []
        for src in srcs:
            obj = f"def{nd}_{src[:-4]}.o"
            run([fc, "-O2", "-c", os.path.join(FORTRAN, "defaults", src), "-o", obj])
            objs.append(obj)
        arc = f"libschooks{nd}.a"
        if os.path.exists(os.path.join(bdir, arc)):
            os.remove(os.path.join(bdir, arc))
        run(["ar", "rcs", arc] + objs)
    out = []
    for app, (nd, srcs) in FAPPS.items():
        objs = []
        for src in srcs:
            obj = f"{app}_{src}.o"
            run([fc, "-O2", "-c", os.path.join(FORTRAN, "apps", src), "-o", obj])
            objs.append(obj)
        exe = f"xsclaw_gpu_{app}"
        run([fc, "-O2", "-o", exe] + drv + objs + [f"-L{bdir}", f"-lschooks{nd}", f"-L{HERE}",
                                                   "-lsharpclaw_amd", "-Wl,-rpath,$ORIGIN/../.."])
        out.append(os.path.join(bdir, exe))
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
