"""Build the gfx950 library, the CLI tools and the synthetic-data helper.

Outputs (git-ignored, shipped with the source tree):
  trinityrnaseq_amd/lib/libtrinity_gpu.so   HIP kernels + C ABI (include/trinity_gpu.h)
  trinityrnaseq_amd/lib/libtk_synth.so      seeded synthetic read generator (bench/tests)
  trinityrnaseq_amd/lib/libtk_micro.so      random-access ceiling microbenchmarks (bench)
  trinityrnaseq_amd/bin/*                   CLI tools keeping the reference command lines
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TK_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-Wall", "-pthread", f"-I{ROOT / 'include'}", f"-I{CSRC}"]

HIP_SOURCES = ["tk_lib.hip", "tk_sort.hip", "tk_batch.hip", "tk_iworm.hip", "tk_kset.hip"]
HOST_SOURCES = ["tk_server.cpp"]
TOOLS = {
    "jellyfish": ["tools/jellyfish_main.cpp"],
    "fastaToKmerCoverageStats": ["tools/fastaToKmerCoverageStats_main.cpp"],
    "inchworm": ["tools/inchworm_main.cpp"],
    "ReadsToTranscripts": ["tools/ReadsToTranscripts_main.cpp"],
    "nbkc_normalize": ["tools/nbkc_normalize_main.cpp"],
    "fastQ_to_fastA": ["tools/fastq_to_fasta_main.cpp"],
    "seqtk-trinity": ["tools/fastq_to_fasta_main.cpp"],
    "extract_normalized_reads": ["tools/extract_normalized_reads_main.cpp"],
    "GraphFromFasta": ["tools/GraphFromFasta_main.cpp"],
    "FastaToDeBruijn": ["tools/FastaToDeBruijn_main.cpp"],
    "QuantifyGraph": ["tools/QuantifyGraph_main.cpp"],
    "insilico_read_normalization": ["tools/insilico_read_normalization_main.cpp"],
}
TOOL_COMMON = ["tools/seq_io.cpp", "tools/tool_util.cpp", "tools/ingest.cpp", "tools/extract.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def _newer(out: Path, deps: list[Path]) -> bool:
    return out.exists() and all(out.stat().st_mtime >= d.stat().st_mtime for d in deps)


def build_library(force: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.h")) + [ROOT / "include" / "trinity_gpu.h"]
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        s = CSRC / src
        o = OBJ / (s.stem + ".o")
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)])
    for src in HOST_SOURCES:
        s = CSRC / src
        o = OBJ / (s.stem + ".o")
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs.append([CXX, "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", f"-I{ROOT / 'include'}",
                         "-c", str(s), "-o", str(o)])
    with ThreadPoolExecutor(max_workers=max(1, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    out = LIB / "libtrinity_gpu.so"
    tmp = LIB / "libtrinity_gpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_synth(force: bool = True) -> Path:
    LIB.mkdir(exist_ok=True)
    src = CSRC / "synth" / "tk_synth.cpp"
    out = LIB / "libtk_synth.so"
    if force or not _newer(out, [src]):
        tmp = LIB / "libtk_synth.so.tmp"
        _run([CXX, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", str(src), "-o", str(tmp)])
        os.replace(tmp, out)
    return out


def build_micro(force: bool = True) -> Path:
    LIB.mkdir(exist_ok=True)
    src = CSRC / "tk_micro.hip"
    out = LIB / "libtk_micro.so"
    if force or not _newer(out, [src]):
        tmp = LIB / "libtk_micro.so.tmp"
        _run([HIPCC, *HIP_FLAGS, "-shared", str(src), "-o", str(tmp)])
        os.replace(tmp, out)
    return out


def build_tools(force: bool = True) -> list[Path]:
    BIN.mkdir(exist_ok=True)
    common = [CSRC / c for c in TOOL_COMMON]
    outs = []
    jobs = []
    for name, srcs in TOOLS.items():
        paths = [CSRC / s for s in srcs] + common
        if not all(p.exists() for p in paths):
            continue
        out = BIN / name
        outs.append(out)
        deps = paths + list((CSRC / "tools").glob("*.h")) + [ROOT / "include" / "trinity_gpu.h"]
        if force or not _newer(out, deps):
            jobs.append([CXX, *CXX_FLAGS, *map(str, paths), "-o", str(out),
                         f"-L{LIB}", "-ltrinity_gpu", "-lz", "-Wl,-rpath,$ORIGIN/../lib"])
    with ThreadPoolExecutor(max_workers=max(1, min(8, len(jobs)))) as ex:
        list(ex.map(_run, jobs))
    return outs


def build_all(force: bool = True) -> None:
    build_library(force)
    build_synth(force)
    build_micro(force)
    build_tools(force)


if __name__ == "__main__":
    build_all(force="--incremental" not in sys.argv)
    print("built", LIB, BIN)
