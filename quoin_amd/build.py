"""Build libquoin_amd.so (gfx950) in-tree with hipcc.

Every HIP source is compiled for gfx950 with FP contraction off (the ALP and
FloatMult checks are bit-exact IEEE FP64) and packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libquoin_amd.so")
SOURCES = ["qa_scan.hip", "qa_encode.hip", "qa_fast.hip", "qa_fast_emit.hip", "qa_decode.hip",
           "qa_fast_dec.hip", "qa_dec_tiles.hip", "qa_rans.hip", "qa_resid.hip", "qa_host.cpp", "qa_server.cpp",
           "qa_entropy.cpp", "qa_decimal.hip", "qa_pco.cpp", "qa_hybrid.cpp", "qa_noalloc.hip", "qa_pool.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-fno-strict-aliasing", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
    "-Wno-unused-result",
]


def _compile(src, bdir, verbose):
    obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msg}")
    if verbose:
        print(" ".join(cmd), flush=True)
        if msg:
            print(msg)
    return obj


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    jobs = min(len(SOURCES), os.cpu_count() or 1, 8)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, bdir, verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
