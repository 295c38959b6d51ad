"""Build the HIP library gpe_amd/libgpehip.so for gfx950 (in-tree).

    python -m gpe_amd.build

Each translation unit is compiled with hipcc; the FP64 one with
-ffp-contract=off so its arithmetic is one rounding per operation in the
reference's order.  Packed FP32 instructions are disabled on every unit.

Every build recompiles every unit, holds a lock on its object folder (two
builds never interleave their objects) and refuses to link when a file under
csrc/ changed while it compiled: the host code and the kernels share
argument structs by layout, and objects compiled from two versions of a
header would link without complaint.
"""
import fcntl
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpehip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]
N_VARIANTS = 5  # kTiles in csrc/launch.h
UNITS = [("gpe_misc.hip", "gpe_misc", ["-ffp-contract=off"]),
         ("gpe_capi.cpp", "gpe_capi", ["-ffp-contract=off", "-x", "hip"]),
         ("gpe_batch.cpp", "gpe_batch", ["-ffp-contract=off", "-x", "hip"]),
         ("host_common.cpp", "host_common", ["-ffp-contract=off", "-x", "hip"]),
         ("gpe_ic.cpp", "gpe_ic", ["-ffp-contract=off", "-x", "hip"]),
         ("tune_table.cpp", "tune_table", ["-x", "hip"]),
         ("gpe_spectral.cpp", "gpe_spectral", ["-ffp-contract=off", "-x", "hip"]),
         ("gpe_noalloc.cpp", "gpe_noalloc", ["-ffp-contract=off", "-x", "hip"]),
         ("gpe_serve.cpp", "gpe_serve", ["-x", "hip"]),
         ("errors.cpp", "errors", ["-x", "hip"]),
         ("ranges.cpp", "ranges", ["-x", "hip"])]
for _v in range(N_VARIANTS):
    # stage outputs are written with nontemporal stores (-DGPE_NT_STORE): they
    # are not re-read within a sweep, and keeping them out of L2 leaves room
    # for the ghost rows neighbouring tiles share (measured 3% faster)
    UNITS.append(("gpe_stage.hip", "stage_f64_v%d" % _v,
                  ["-ffp-contract=off", "-DGPE_F64", "-DGPE_VID=%d" % _v, "-DGPE_NT_STORE"]))
    # FP32 too is built without contraction: each operation rounds once, so
    # the single-simulation and the batched instantiation of a stage (inlined
    # into different kernels) compute bitwise the same field
    UNITS.append(("gpe_stage.hip", "stage_f32_v%d" % _v, ["-ffp-contract=off", "-DGPE_VID=%d" % _v, "-DGPE_NT_STORE"]))

# fused stage pairs: FP32 forms only (three tile variants)
for _v in range(3):
    UNITS.append(("gpe_pair.hip", "pair_f32_v%d" % _v, ["-ffp-contract=off", "-DGPE_PVID=%d" % _v, "-DGPE_NT_STORE"]))


def _compile(unit):
    src, name, extra = unit
    out = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return out


def _sources_digest():
    h = hashlib.sha256()
    for name in sorted(os.listdir(CSRC)):
        path = os.path.join(CSRC, name)
        if os.path.isfile(path):
            h.update(name.encode() + b"\0")
            with open(path, "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()


def build(verbose=True, extra_flags=(), lib_path=None, obj_dir=None):
    """Compile every unit and link the library.  extra_flags/lib_path/obj_dir
    build an alternative library (e.g. -DGPE_NT_STORE) beside the default."""
    global OBJ
    if obj_dir:
        OBJ = obj_dir
    os.makedirs(OBJ, exist_ok=True)
    if extra_flags:
        global UNITS
        UNITS = [(a, b, c + list(extra_flags)) for a, b, c in UNITS]
    with open(os.path.join(OBJ, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        before = _sources_digest()
        with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 4)) as ex:
            objs = list(ex.map(_compile, UNITS))
        if _sources_digest() != before:
            raise RuntimeError("files under %s changed during the build: its objects may mix two versions of "
                               "the shared headers; build again" % CSRC)
        return _link(objs, verbose, lib_path)


def _link(objs, verbose, lib_path):
    out = lib_path or LIB
    tmp = "%s.tmp%d" % (out, os.getpid())
    # hipFFT (the random-phase initial state) from the ROCm install
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib")
    cmd = ([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
           ["-L" + rocm_lib, "-lhipfft", "-lrocprofiler-sdk-roctx", "-Wl,-rpath," + rocm_lib])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
