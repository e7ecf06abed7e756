"""Build libsolhip.so (HIP, gfx950) in-tree.

    python sol-rs_amd/build.py            # compile every csrc/*.hip, link lib/libsolhip.so,
                                          # then the host mirror, examples and bin/valu_peak

Objects go to sol-rs_amd/_build/, the libraries to sol-rs_amd/lib/ and the
example programs to sol-rs_amd/bin/.  hipcc cross-compiles for gfx950 without
a GPU; the host mirror (host/*.cpp, the C++ face of the sol crate over the C
ABI) and the examples are plain g++.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsolhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HOST = os.path.join(HERE, "host")
EXAMPLES = os.path.join(HERE, "examples")
BINDIR = os.path.join(HERE, "bin")
HOST_LIB = os.path.join(LIBDIR, "libsolrs.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-I" + INCLUDE]

# -ffp-contract=off: the fp32 arithmetic is evaluated exactly as written (the
#   parity contract with the host restatement);
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 generation.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src, build_dir=BUILD, defines=()):
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-D" + d for d in defines] + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=True, out=LIB, build_dir=BUILD, defines=()):
    """compile every csrc/*.hip (optionally with -D defines into another
    build directory / library path) and link the shared library"""
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, warn in ex.map(lambda s: _compile(s, build_dir, defines), srcs):
            objs.append(obj)
            if warn and verbose:
                print(warn, file=sys.stderr)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s\n%s" % (cmd[0], " ".join(cmd), r.stderr))


def build_host(verbose=True):
    """lib/libsolrs.so (host/*.cpp, linked against libsolhip.so and zlib) and
    one program per examples/*.cpp in bin/, found at run time through an
    $ORIGIN-relative rpath"""
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    tmp = HOST_LIB + ".tmp"
    _run([CXX] + CXX_FLAGS + ["-shared", "-o", tmp] + srcs +
         ["-L" + LIBDIR, "-lsolhip", "-lz", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    progs = []
    for src in sorted(glob.glob(os.path.join(EXAMPLES, "*.cpp"))):
        exe = os.path.join(BINDIR, os.path.splitext(os.path.basename(src))[0])
        _run([CXX] + CXX_FLAGS + ["-o", exe + ".tmp", src, "-L" + LIBDIR, "-lsolrs", "-lsolhip",
                                  "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(exe + ".tmp", exe)
        progs.append(exe)
    if verbose:
        print("built", HOST_LIB, "and", ", ".join(os.path.basename(p) for p in progs))
    return HOST_LIB


def build_measure(verbose=True):
    """bin/valu_peak: the fp32 VALU issue-rate probe (measure/valu_peak.hip)
    whose figure the traversal kernels' counters are compared with"""
    os.makedirs(BINDIR, exist_ok=True)
    src = os.path.join(HERE, "measure", "valu_peak.hip")
    exe = os.path.join(BINDIR, "valu_peak")
    _run([HIPCC] + HIP_FLAGS + ["-o", exe + ".tmp", src])
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    build()
    build_host()
    build_measure()
