"""Native build of the package: the HIP C-ABI library (libmck.so, gfx950), the synthetic
workload generator (libmck_synth.so) and, as test infrastructure, the CPU oracle
(oracle/liboracle.so) and the reference's own kdtree2 (oracle/_ref)."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# Packed FP32 instructions are disabled for every HIP translation unit; FMA contraction
# is disabled so that the point-location arithmetic matches the reference bit for bit.
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def _stale(target, sources):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in sources)


def build_synth(force=True):
    # no FMA contraction: the wavelet arithmetic (synth_math.h) must round as the device
    # generator's does
    src = os.path.join(CSRC, "synth.cpp")
    out = os.path.join(HERE, "libmck_synth.so")
    if force or _stale(out, [src, os.path.join(CSRC, "synth_math.h")]):
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", src, "-o", out])
    return out


def hip_sources():
    return [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))
            if f.endswith(".hip") and not f.startswith("synth")]


def build_synth_gpu(force=True):
    """Device-side generator of synthetic databases (benchmark input, not the product)."""
    src = os.path.join(CSRC, "synth_gpu.hip")
    out = os.path.join(HERE, "libmck_synth_gpu.so")
    if force or _stale(out, [src, os.path.join(CSRC, "synth_math.h")]):
        hipcc = os.path.join(ROCM, "bin", "hipcc")
        _run([hipcc] + HIP_FLAGS + ["-shared", src, "-o", out])
    return out


def build_hip(force=True):
    srcs = hip_sources()
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs.append(os.path.join(ROOT, "include", "mck.h"))
    out = os.path.join(HERE, "libmck.so")
    if not srcs:
        raise RuntimeError("no HIP sources")
    if force or _stale(out, srcs + hdrs):
        hipcc = os.path.join(ROCM, "bin", "hipcc")
        objs = []
        for s in srcs:
            o = os.path.join(HERE, "_build", os.path.basename(s) + ".o")
            os.makedirs(os.path.dirname(o), exist_ok=True)
            _run([hipcc] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                        "-c", s, "-o", o])
            objs.append(o)
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs +
             ["-o", out, "-Wl,--no-undefined", "-lamdhip64"])
    return out


def build_oracle(force=True):
    odir = os.path.join(ROOT, "oracle")
    srcs = [os.path.join(odir, f) for f in ("mck_oracle_core.c", "mck_oracle_path.c", "mck_oracle_results.c")]
    out = os.path.join(odir, "liboracle.so")
    if force or _stale(out, srcs + [os.path.join(odir, "mck_oracle.h")]):
        _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off",
              "-fno-fast-math", "-fexcess-precision=standard"] + srcs + ["-o", out, "-lm"])
    return out


def build_ref(force=True):
    """oracle/_ref: the reference's kdtree2.f90 compiled where it lies, plus this repo's
    bind(c) wrapper; skipped when the reference tree is absent."""
    odir = os.path.join(ROOT, "oracle")
    script = os.path.join(odir, "build_ref.sh")
    if not os.path.isdir("/root/reference/src") or not shutil.which("bash"):
        return None
    out = os.path.join(odir, "_ref", "libref_kdtree.so")
    if force or not os.path.exists(out):
        _run(["bash", script])
    return out


HDF5_PREFIX = os.environ.get("MCK_HDF5_PREFIX", "/opt/conda")


def build_dbio(force=True):
    """libmck_dbio.so: merged database files through HDF5 (host code, g++); skipped when
    the HDF5 headers and library are absent."""
    src = os.path.join(CSRC, "mck_dbio.cpp")
    out = os.path.join(HERE, "libmck_dbio.so")
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(lib, "libhdf5.so"))):
        print("libmck_dbio.so skipped: no HDF5 under %s" % HDF5_PREFIX, file=sys.stderr)
        return None
    hdrs = [os.path.join(ROOT, "include", h) for h in ("mck_dbio.h", "mck_output.h", "mck.h")]
    if force or _stale(out, [src] + hdrs):
        # libhdf5 by path with an rpath for it; libstdc++ linked statically so that the
        # older C++ runtime beside libhdf5 is never picked up in its place
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
              "-I", inc, src, "-o", out, os.path.join(lib, "libhdf5.so"), "-Wl,-rpath," + lib,
              "-static-libstdc++", "-static-libgcc", "-Wl,--no-undefined"])
    return out


def flang():
    return shutil.which("flang") or shutil.which("amdflang") or (
        os.path.join(ROCM, "llvm", "bin", "amdflang") if os.path.exists(os.path.join(ROCM, "llvm", "bin", "amdflang"))
        else None)


def build_fortran_e2e(force=True):
    """tests/fortran/slave_e2e: a Fortran program that drives libmck.so through the
    iso_c_binding module (mc_kernel_amd/fortran/mck_gpu.f90), test infrastructure of the
    Fortran binding; skipped when no Fortran compiler is found."""
    fc = flang()
    if fc is None:
        print("slave_e2e skipped: no Fortran compiler", file=sys.stderr)
        return None
    mod = os.path.join(HERE, "fortran", "mck_gpu.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "slave_e2e.f90")
    out = os.path.join(ROOT, "tests", "fortran", "slave_e2e")
    if force or _stale(out, [mod, prog]):
        bdir = os.path.join(HERE, "_build", "fortran")
        os.makedirs(bdir, exist_ok=True)
        obj = os.path.join(bdir, "mck_gpu.o")
        _run([fc, "-O1", "-c", mod, "-o", obj, "-module-dir", bdir])
        _run([fc, "-O1", prog, obj, "-I", bdir, "-L" + HERE, "-lmck", "-Wl,-rpath,$ORIGIN/../../mc_kernel_amd",
              "-o", out])
    return out


# host code only (gcc / g++, no HIP source on these lines)
SANITIZE = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=undefined"]


def build_sanitized(outdir):
    """The oracle and libmck_dbio.so built with AddressSanitizer and
    UndefinedBehaviorSanitizer into outdir (test infrastructure); returns their paths
    (the second None without HDF5)."""
    os.makedirs(outdir, exist_ok=True)
    odir = os.path.join(ROOT, "oracle")
    osrc = [os.path.join(odir, f) for f in ("mck_oracle_core.c", "mck_oracle_path.c", "mck_oracle_results.c")]
    ora = os.path.join(outdir, "liboracle_san.so")
    _run(["gcc"] + SANITIZE + ["-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                               "-fexcess-precision=standard"] + osrc + ["-o", ora, "-lm"])
    dbio = None
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if os.path.exists(os.path.join(inc, "hdf5.h")):
        dbio = os.path.join(outdir, "libmck_dbio_san.so")
        _run(["g++"] + SANITIZE + ["-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"), "-I", inc,
                                   os.path.join(CSRC, "mck_dbio.cpp"), "-o", dbio, os.path.join(lib, "libhdf5.so"),
                                   "-Wl,-rpath," + lib])
    return ora, dbio


def build_all(force=True):
    build_synth(force)
    build_oracle(force)
    try:
        build_ref(force)
    except subprocess.CalledProcessError as e:  # the reference build is optional
        print("oracle/_ref build failed:", e, file=sys.stderr)
    build_hip(force)
    build_synth_gpu(force)
    build_dbio(force)
    build_fortran_e2e(force)


if __name__ == "__main__":
    build_all(force="--force" in sys.argv)
