"""Builds libesthera.so (gfx950) in-tree and, for tests, the CPU oracle.

    python -m esthera_amd.build            # engine + oracle (+ reference MTGP where present)

Steps:
  1. esthera_amd/data/mtgp32dc_11213.txt -> _build/gen/mtgp32dc_11213.inc
     (the 200 MTGP32-11213 parameter sets as a C initializer list)
  2. hipcc --offload-arch=gfx950 on every csrc/*.hip, linked into
     esthera_amd/libesthera.so
  3. gcc on csrc/esth_pf_main.c -> esthera_amd/esthera_pf (the command
     line, C host code over the C ABI)
  4. make -C oracle (test infrastructure), and make -C oracle ref when the
     reference tree is present (not needed where only the built library runs).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(REPO, "_build")
GEN = os.path.join(BUILD, "gen")
LIB = os.path.join(HERE, "libesthera.so")
ARCH = os.environ.get("ESTH_ARCH", "gfx950")

HIP_SOURCES = ["esth_engine.hip", "esth_filter.hip", "esth_batch.hip", "esth_server.hip"]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every float operation is the one written in the source: host and device agree bit for bit
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-unused-result",
    "-Wno-unused-value",
    f"-I{os.path.join(REPO, 'include')}",
    f"-I{CSRC}",
    f"-I{GEN}",
]


def gen_param_inc():
    os.makedirs(GEN, exist_ok=True)
    src = os.path.join(HERE, "data", "mtgp32dc_11213.txt")
    rows = []
    with open(src) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            t = line.split()
            words = [int(t[0]), int(t[1]), int(t[2])] + [int(x, 16) for x in t[3:]]
            if len(words) != 52:
                raise ValueError("bad parameter row")
            rows.append(", ".join(f"0x{w:08x}u" for w in words))
    if len(rows) != 200:
        raise ValueError(f"expected 200 parameter sets, found {len(rows)}")
    out = os.path.join(GEN, "mtgp32dc_11213.inc")
    text = "/* generated from esthera_amd/data/mtgp32dc_11213.txt */\n" + ",\n".join(rows) + "\n"
    if not os.path.exists(out) or open(out).read() != text:
        with open(out, "w") as fh:
            fh.write(text)


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_engine():
    gen_param_inc()
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


CLI = os.path.join(HERE, "esthera_pf")


def build_cli():
    """esthera_pf: the reference's particle_filter command line over the C ABI
    (plain C, as the reference's driver; linked against libesthera.so next to it)."""
    tmp = CLI + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", f"-I{os.path.join(REPO, 'include')}",
           os.path.join(CSRC, "esth_pf_main.c"), f"-L{HERE}", "-lesthera", "-lm", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"esthera_pf build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)
    return CLI


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if os.path.isdir("/root/reference/src/partfilt/opencl/mtgp"):
        r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference MTGP build failed:\n{r.stdout}\n{r.stderr}")


def main():
    build_engine()
    build_cli()
    build_oracle()
    print(LIB)


if __name__ == "__main__":
    sys.exit(main())
