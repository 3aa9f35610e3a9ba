"""Build libgpr_mi355x.so (HIP, gfx950) in-tree.

    python -m phd_amd.build            # or phd_amd.build.build()

Compiles phd_amd/csrc/*.hip with hipcc for gfx950 only, each unit (and each
variant of a unit compiled more than once) as its own object, in parallel.
Packed FP32 VALU instructions are disabled (-packed-fp32-ops) for every
device compile.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, os.environ.get("GPR_LIB_NAME", "libgpr_mi355x.so"))
ARCH = os.environ.get("GPR_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    # fuse a*b+c into an FMA only within one source expression, so that every
    # kernel instantiation of the same code rounds alike (the one- and
    # two-launch predictors, the batched and single-domain steps)
    "-ffp-contract=on",
] + os.environ.get("GPR_EXTRA_FLAGS", "").split()


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# units compiled more than once, with these macro sets (one object each)
VARIANTS = {
    "gpr_adj_heavy.hip": [{"GPR_ADJ_DUAL": d, "GPR_ADJ_TH": t} for d in (0, 1) for t in (0, 1)],
    "gpr_pred.hip": [{"GPR_UNIT_N": n} for n in (2, 3)],
    "gpr_fv.hip": [{"GPR_UNIT_N": n, "GPR_UNIT_ND": d} for n in (2, 3) for d in (1, 2)],
}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def units():
    """(source, object name, extra flags) for every object of the library"""
    out = []
    for src in sources():
        base = os.path.basename(src)
        for defs in VARIANTS.get(base, [{}]):
            tag = "".join(f"_{k.split('_')[-1].lower()}{v}" for k, v in defs.items())
            out.append((src, base + tag + ".o", [f"-D{k}={v}" for k, v in defs.items()]))
    return out


def build(verbose=False, jobs=None):
    objs = []
    procs = []
    bdir = os.path.join(HERE, "_build", os.path.splitext(os.path.basename(LIB))[0])
    os.makedirs(bdir, exist_ok=True)
    for src, oname, extra in units():
        obj = os.path.join(bdir, oname)
        cmd = [hipcc()] + HIPCC_FLAGS + extra + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    for p, src in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {src}")
        elif verbose and out:
            sys.stderr.write(out.decode(errors="replace"))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_flopcount()
    return LIB


FLOPCOUNT = os.path.join(HERE, "flopcount")


def build_flopcount():
    """The host program that counts the kernels' FP64 operations
    (csrc/flopcount.cpp, the physics headers compiled for the host)."""
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("no host C++ compiler for flopcount")
    subprocess.check_call([cxx, "-std=c++17", "-O1", "-DGPR_HOST_ONLY", "-Wno-unknown-pragmas",
                           os.path.join(CSRC, "flopcount.cpp"), "-o", FLOPCOUNT])
    return FLOPCOUNT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
