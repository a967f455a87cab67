"""In-tree build of the native libraries (no JIT cache, no install step).

* graphaligner_amd/_build/libgraphaligner_amd.so: the product (host C++ + gfx950 HIP)
* graphaligner_amd/_build/libga_synth.so: synthetic workload generator
* oracle/_build/libga_oracle.so and tests/_build/libga_emu.so: test infrastructure
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "graphaligner_amd")
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-variable", "-Wno-sign-compare", "-Wno-unused-function"]
PRODUCT_SOURCES = ["ga_kernels.hip", "ga_mlp.hip", "ga_index_gpu.hip", "ga_runtime.cpp", "ga_graph.cpp", "ga_align.cpp", "ga_proposal.cpp", "ga_index.cpp",
                   "ga_capi.cpp"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout[-8000:])
    return r.stdout


def build_product():
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in PRODUCT_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = os.path.join(OUT, "libgraphaligner_amd.so.tmp")
    _run(["hipcc"] + HIP_FLAGS + ["-o", tmp] + srcs + ["-lpthread"])
    os.replace(tmp, os.path.join(OUT, "libgraphaligner_amd.so"))


def build_calibration():
    """Copy kernel of known size for calibrating the HBM traffic counters."""
    os.makedirs(OUT, exist_ok=True)
    _run(["hipcc"] + HIP_FLAGS + ["-o", os.path.join(OUT, "libga_calib.so"), os.path.join(CSRC, "ga_calib.hip")])


def build_synth():
    os.makedirs(OUT, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", os.path.join(OUT, "libga_synth.so"),
          os.path.join(CSRC, "ga_synth.cpp")])


def build_test_infrastructure():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    tb = os.path.join(ROOT, "tests", "_build")
    os.makedirs(tb, exist_ok=True)
    common = ["g++", "-O2", "-g", "-std=c++17", "-fPIC", "-shared", "-Wno-unused-variable"]
    srcs = [os.path.join(ROOT, "tests", "emu", "ga_emu.cpp"), os.path.join(CSRC, "ga_graph.cpp")]
    # device extension code built for the host: lane helpers as 64-lane loops, and
    # the one-extension-per-thread helpers the gfx950 kernel uses
    _run(common + ["-o", os.path.join(tb, "libga_emu.so")] + srcs)
    _run(common + ["-DGA_THREAD_PER_EXTENSION", "-DGA_THREAD_HOST_TEST", "-o", os.path.join(tb, "libga_thread.so")] + srcs)
    # the wave-cooperative lane code on 64 host threads per wave that meet at a
    # barrier in every wave intrinsic (divergent lanes are reported, not hung)
    _run(common + ["-pthread", "-o", os.path.join(tb, "libga_simt.so"), os.path.join(ROOT, "tests", "emu", "ga_simt.cpp"),
                   os.path.join(CSRC, "ga_graph.cpp")])
    # the product's host pipeline over a backend running the device code on host threads
    pipe = [os.path.join(ROOT, "tests", "emu", "ga_pipeline_emu.cpp")] + [os.path.join(CSRC, s) for s in ("ga_align.cpp", "ga_proposal.cpp", "ga_index.cpp", "ga_graph.cpp")]
    _run(common + ["-pthread", "-o", os.path.join(tb, "libga_pipe.so")] + pipe)
    # the same with the traceback recomputing every column it enters (all of the
    # reference's recalcNodeWordslice checks)
    _run(common + ["-pthread", "-DGA_FULL_RECALC_CHECKS", "-o", os.path.join(tb, "libga_pipe_fullcheck.so")] + pipe)


def build_all():
    build_synth()
    build_calibration()
    build_test_infrastructure()
    build_product()


if __name__ == "__main__":
    build_all()
