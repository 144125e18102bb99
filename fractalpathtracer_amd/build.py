"""Build recipe for libfpt_hip.so (hipcc, gfx950).

Every HIP source under csrc/ is compiled separately (in parallel) and linked
into fractalpathtracer_amd/libfpt_hip.so.  Packed FP32 instructions are
disabled (`-target-feature -packed-fp32-ops`): the compiler would otherwise
emit v_pk_* FP32 forms on its own.  Contraction is `-ffp-contract=on`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libfpt_hip.so"
ARCH = os.environ.get("FPT_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # FMA fusion only within one source expression: results then do not
    # depend on where the optimiser inlines a function (files that must be
    # IEEE-exact switch contraction off with a pragma).
    "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(CSRC.glob("*.hip"))


# Device headers the runtime compiler needs for topology-specialised kernels,
# under the include names the sources use.
EMBEDDED_HEADERS = [
    ("fpt_device.h", CSRC / "fpt_device.h"),
    ("fpt_internal.h", CSRC / "fpt_internal.h"),
    ("learned.h", CSRC / "learned.h"),
    ("persistent_kernel.h", CSRC / "persistent_kernel.h"),
    ("../../include/fpt_abi.h", PKG.parent / "include" / "fpt_abi.h"),
]


def write_embedded_headers() -> Path:
    """_build/embedded_headers.h: the headers as string constants."""
    out = ["// Generated by build.py: device headers embedded for the runtime compiler.", "#pragma once",
           "#include <stddef.h>", "namespace fpt {"]
    names, texts = [], []
    for i, (name, path) in enumerate(EMBEDDED_HEADERS):
        text = path.read_text()
        # split into raw-literal chunks (compiler string-literal limits)
        chunks = [text[j:j + 16000] for j in range(0, len(text), 16000)] or [""]
        lit = "\n".join(f'R"FPTHDR({c})FPTHDR"' for c in chunks)
        out.append(f"static const char kEmbeddedHeader{i}[] = {lit};")
        names.append(f'"{name}"')
        texts.append(f"kEmbeddedHeader{i}")
    out.append(f"static const size_t kEmbeddedHeaderCount = {len(EMBEDDED_HEADERS)};")
    out.append("static const char *kEmbeddedHeaderNames[] = {" + ", ".join(names) + "};")
    out.append("static const char *kEmbeddedHeaderText[] = {" + ", ".join(texts) + "};")
    out.append("}  // namespace fpt")
    path = BUILD / "embedded_headers.h"
    new = "\n".join(out) + "\n"
    if not path.exists() or path.read_text() != new:
        path.write_text(new)
    return path


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every csrc/*.hip for gfx950 and link libfpt_hip.so."""
    BUILD.mkdir(exist_ok=True)
    write_embedded_headers()
    cc = hipcc()
    srcs = sources()

    def compile_one(src: Path) -> Path:
        obj = BUILD / (src.stem + ".o")
        cmd = [cc, *COMMON, "-I", str(BUILD), "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-8000:]}")
        return obj

    n = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(max_workers=n) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lz", "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
