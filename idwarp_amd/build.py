"""In-tree builds: the HIP library (gfx950) and, for tests, a HIP test library
and the CPU oracle."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["idw_api.hip", "idw_kernels.hip", "idw_tree.cpp"]
LIB = os.path.join(HERE, "libidwarp_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-munsafe-fp-atomics",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: the fast-mode loops carry them across
    # iterations, and the AGPR form copied them back and forth every step
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build_library():
    """Compile libidwarp_hip.so with hipcc for gfx950 (always rebuilds)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    _run([hipcc] + HIPCC_FLAGS + ["-o", tmp] + SOURCES, CSRC)
    os.replace(tmp, LIB)
    return LIB


TEST_LIB = os.path.join(ROOT, "tests", "native", "libwalkmath.so")


def build_test_library():
    """Compile tests/native/libwalkmath.so, the walk's device arithmetic on
    arrays for the tests (same flags as the library)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = TEST_LIB + ".tmp"
    _run([hipcc] + HIPCC_FLAGS + ["-I", CSRC, "-o", tmp, "walkmath.hip"], os.path.dirname(TEST_LIB))
    os.replace(tmp, TEST_LIB)
    return TEST_LIB


HOST_CHECKS = os.path.join(ROOT, "tests", "native", "host_checks")


def build_host_checks():
    """Compile tests/native/host_checks: the library's host C++ (the tree
    build, the C-ABI's host side) with AddressSanitizer and
    UndefinedBehaviorSanitizer, for CPU tests.  Device code is compiled
    without them (-Xarch_host); the objects are linked by clang++, whose
    sanitizer runtime the HIP objects need."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    clang = os.path.join(os.path.dirname(os.path.realpath(hipcc)), "..", "llvm", "bin", "clang++")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang++"
    out = os.path.join(ROOT, "tests", "native", "_build")
    os.makedirs(out, exist_ok=True)
    dev = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    host_san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                "-Xarch_host", "-fno-sanitize-recover=undefined"]
    cpp = [clang, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    obj = {k: os.path.join(out, k + ".o") for k in ("kernels", "api", "tree", "checks")}
    _run([hipcc] + dev + ["-c", "-o", obj["kernels"], os.path.join(CSRC, "idw_kernels.hip")], CSRC)
    _run([hipcc] + dev + host_san + ["-c", "-o", obj["api"], os.path.join(CSRC, "idw_api.hip")], CSRC)
    _run(cpp + ["-c", "-o", obj["tree"], os.path.join(CSRC, "idw_tree.cpp")], CSRC)
    _run(cpp + ["-c", "-o", obj["checks"], os.path.join(ROOT, "tests", "native", "host_checks.cpp")], CSRC)
    rocm_lib = os.path.join(os.path.dirname(os.path.realpath(hipcc)), "..", "lib")
    tmp = HOST_CHECKS + ".tmp"
    _run([clang, "-fsanitize=address,undefined", "-o", tmp, obj["checks"], obj["api"], obj["tree"], obj["kernels"],
          "-L" + rocm_lib, "-lamdhip64", "-Wl,-rpath," + rocm_lib], CSRC)
    os.replace(tmp, HOST_CHECKS)
    return HOST_CHECKS


def build_oracle(reference="/root/reference/src"):
    """Compile oracle/liboracle.so and, where the reference sources exist,
    oracle/_ref/libidwref.so (test-only checkers)."""
    targets = ["all"]
    if os.path.isdir(reference):
        targets.append("ref")
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "REF=" + reference] + targets, ROOT)


if __name__ == "__main__":
    print(build_library())
    print(build_test_library())
    print(build_host_checks())
    build_oracle()
