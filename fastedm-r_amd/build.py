"""Builds libfastedm_hip.so in-tree for gfx950 with hipcc (no JIT cache: the library travels with the tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfastedm_hip.so")
SOURCES = ["edm_api.hip", "kernels_plain.hip", "kernels_batch.hip", "kernels_rank.hip", "kernels_mfma.hip", "kernels_select.hip", "kernels_wasserstein.hip", "kernels_smap.hip", "edm_host.cpp"]
# -ffp-contract=off: FP64 products and sums are rounded one by one, as on the CPU path.
# packed-fp32-ops disabled: no v_pk_* FP32 instructions are generated.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result", "-Wno-unused-value"]


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", src + ".o")
        cmd = ["hipcc", "-c"] + [f for f in FLAGS if f != "-shared"] + ["-I", os.path.join(HERE, "..", "include"),
                                                                       os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = ["hipcc", "-c", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I",
                   os.path.join(HERE, "..", "include"), os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    ok = True
    for cmd, p in procs:
        out = p.communicate()[0].decode()
        if p.returncode != 0:
            ok = False
            print(" ".join(cmd), "\n", out, file=sys.stderr)
        elif verbose and out.strip():
            print(out)
    if not ok:
        raise RuntimeError("hipcc failed")
    link = ["hipcc", "-shared", "-fPIC", "--offload-arch=gfx950", "-o", LIB + ".tmp"] + objs + ["-lpthread", "-Wl,--no-undefined"]
    subprocess.run(link, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
