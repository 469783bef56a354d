"""Build the HIP backend in-tree.

Products (all git-ignored, all shipped with the source tree):
  nemo_amd/lib/libnemo_hip.so              backend + C ABI (include/nemo_hip.h)
  nemo_amd/plugins/hip/lib<Model>.so       one neuron-model plugin per model

Device code is compiled for gfx950 only, with FP contraction off (the CPU
reference is evaluated without fused multiply-add), correctly rounded fp32
division and square root, denormals preserved, and the packed-fp32 target
feature disabled.  Host code is compiled by g++ with OpenMP (connectivity
builder).

Usage: python -m nemo_amd.build  (or nemo_amd.build.build())
"""
import concurrent.futures
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nemo_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
PLUGIN_OUT = os.path.join(PKG, "plugins", "hip")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

MODELS = ["Izhikevich", "QIF", "IF_curr_exp", "HH", "PoissonSource", "Input", "Kuramoto", "KuramotoLag",
          "IzhikevichChaos", "IzhikevichSyn", "Izhikevich2007", "Izhikevich2007DopSyn", "BayesianNeuron",
          "InhomogeneousPoissonSource"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
CXX_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
    "-Wall", "-Wno-unused-result", "-Wno-deprecated-declarations",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-I" + os.path.join(ROCM, "include"),
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    return proc.stdout


def _hipcc():
    return os.path.join(ROCM, "bin", "hipcc")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(PLUGIN_OUT, exist_ok=True)
    jobs = []
    # backend objects
    for src in ("kernels.hip", "float_mode.hip"):
        jobs.append([_hipcc()] + HIP_FLAGS + ["-c", os.path.join(CSRC, src),
                                              "-o", os.path.join(BUILD, src.replace(".hip", ".o"))])
    for src in ("backend.cpp", "float_network.cpp", "nemo_c.cpp"):
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src),
                                           "-o", os.path.join(BUILD, src.replace(".cpp", ".o"))])
    # plugins: one shared library per model
    for m in MODELS:
        jobs.append([_hipcc()] + HIP_FLAGS + ["-shared", os.path.join(CSRC, "plugins", m + ".hip"),
                                              "-o", os.path.join(PLUGIN_OUT, "lib%s.so" % m)])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    link = ["g++", "-shared", "-o", os.path.join(LIBDIR, "libnemo_hip.so"),
            os.path.join(BUILD, "kernels.o"), os.path.join(BUILD, "float_mode.o"),
            os.path.join(BUILD, "backend.o"), os.path.join(BUILD, "float_network.o"),
            os.path.join(BUILD, "nemo_c.o"),
            "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-fopenmp", "-ldl",
            # calls between the library's own entry points bind inside it, even
            # where a process also holds NeMo's libnemo (same nemo_* names)
            "-Wl,-Bsymbolic",
            "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    _run(link)
    return os.path.join(LIBDIR, "libnemo_hip.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
