"""Builds the package's native library ``libpsst.so`` in-tree with hipcc for
gfx950: the host C++ (Fluent reader, grid flattening, settings, writers,
generator), the HIP solver kernels behind the C-ABI declared in
``include/psst_gpu.h`` and the warm-start MLP kernels of ``include/psst_mlp.h``."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsst.so")

HOST_SOURCES = ["mesh_io.cpp", "grid.cpp", "setup.cpp", "output.cpp", "meshgen.cpp", "host_api.cpp", "tiles.cpp",
                "ordered_sched.cpp"]
DEVICE_SOURCES = ["psst_device.hip", "mlp.hip"]
# the solver's launch templates, instantiated once per (dimension, state type)
INSTANCES = [(2, "double"), (2, "float"), (3, "double"), (3, "float")]

# -ffp-contract=off: every multiply and add is rounded separately, as in the
# x86-64 reference build (no FMA), which FP64 bitwise parity requires.
# -packed-fp32-ops off: no v_pk_*_f32 instructions.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    """Compiles every object (in parallel) and links libpsst.so."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src.replace(".cpp", ".o"))
        cmds.append([cc, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON)
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmds.append([cc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE)
        objs.append(obj)
    for dim, t in INSTANCES:
        obj = os.path.join(objdir, "psst_inst_%dd_%s.o" % (dim, t))
        cmds.append([cc, "-x", "hip", "-c", os.path.join(CSRC, "psst_inst.hip"), "-o", obj,
                     "-DPSST_INST_DIM=%d" % dim, "-DPSST_INST_T=%s" % t] + COMMON + DEVICE)
        objs.append(obj)
    # the slow device objects first
    order = sorted(range(len(cmds)), key=lambda i: "-x" in cmds[i] and cmds[i][2] == "hip", reverse=True)
    jobs = jobs or max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda i: _run(cmds[i], verbose), order))
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950"], verbose)
    os.replace(tmp, LIB)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    elif verbose and r.stdout.strip():
        print(r.stdout)


if __name__ == "__main__":
    print(build(verbose=True))
