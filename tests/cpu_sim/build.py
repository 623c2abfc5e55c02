"""Builds the CPU build of libqo (test infrastructure, runs without a GPU).

Every ``query_on_opencl_fpga_amd/csrc`` source is rewritten so that plain
clang++ compiles it for the host: a launch ``k<<<grid, block, lds, stream>>>(args)``
becomes ``::simt::launch(::simt::cfg(grid, block, lds, stream), "k", [=]() { k(args); })``
and ``extern __shared__ T s[]`` becomes a pointer into the simulated LDS.
``hip/hip_runtime.h`` in this folder supplies the execution model (fibers,
wave collectives, a synchronous runtime over host memory).  The result is
``_build/<variant>/libqo_cpu.a`` plus any driver program named on the command
line; SIMT_VARIANT picks asan (default), msan or fast.

    python tests/cpu_sim/build.py            # library only
    python tests/cpu_sim/build.py serve_mix  # library + tests/cpu_sim/serve_mix.cpp
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
CSRC = ROOT / "query_on_opencl_fpga_amd" / "csrc"
CXX = os.environ.get("SIMT_CXX", "/opt/rocm/llvm/bin/clang++")
# Variants: "asan" (AddressSanitizer + UBSan, the default), "msan"
# (MemorySanitizer: reads of device memory nothing wrote), "fast" (none).
VARIANT = os.environ.get("SIMT_VARIANT", "asan")
SAN_FLAGS = {
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-sanitize=vptr"],
    "msan": ["-fsanitize=memory", "-fsanitize-memory-track-origins", "-fsanitize-recover=memory"],
    "fast": ["-O2", "-fPIC"],
}[VARIANT]
OUT = HERE / "_build" / VARIANT
SRC = HERE / "_build" / "src"
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-Wno-unused-result", "-Wno-deprecated",
         f"-I{HERE}", f"-I{SRC}", f"-I{ROOT / 'include'}", *SAN_FLAGS]


def _match_back(s: str, i: int, open_c: str, close_c: str) -> int:
    """Index of the bracket that opens the one closing at s[i]."""
    depth = 0
    while i >= 0:
        if s[i] == close_c:
            depth += 1
        elif s[i] == open_c:
            depth -= 1
            if depth == 0:
                return i
        i -= 1
    raise ValueError("unbalanced brackets")


def _match_fwd(s: str, i: int, open_c: str, close_c: str) -> int:
    depth = 0
    while i < len(s):
        if s[i] == open_c:
            depth += 1
        elif s[i] == close_c:
            depth -= 1
            if depth == 0:
                return i
        i += 1
    raise ValueError("unbalanced brackets")


def rewrite_launches(text: str) -> str:
    out = []
    pos = 0
    while True:
        i = text.find("<<<", pos)
        if i < 0:
            out.append(text[pos:])
            break
        # Kernel expression: identifier (with ::) and optional template arguments.
        j = i - 1
        while text[j].isspace():
            j -= 1
        if text[j] == ">":
            j = _match_back(text, j, "<", ">") - 1
        while j >= 0 and (text[j].isalnum() or text[j] in "_:"):
            j -= 1
        start = j + 1
        name = text[start:i].strip()
        k = text.find(">>>", i)
        cfg = text[i + 3:k]
        p = k + 3
        while text[p].isspace():
            p += 1
        assert text[p] == "(", f"launch of {name} without arguments"
        e = _match_fwd(text, p, "(", ")")
        args = text[p + 1:e]
        base = re.sub(r"<.*", "", name)
        out.append(text[pos:start])
        out.append(f'::simt::launch(::simt::cfg({cfg}), "{base}", [=]() {{ {name}({args}); }})')
        pos = e + 1
    return "".join(out)


EXTERN_LDS = re.compile(r"extern\s+__shared__\s+(?:__attribute__\(\(aligned\(\d+\)\)\)\s+)?([\w:]+(?:\s+[\w:]+)*?)\s+(\w+)\[\];")


CARVE_OLD = """        off += count * sizeof(T);
        return p;"""
CARVE_NEW = """        off += count * sizeof(T);
        if (::simt::canary_on()) {
            if (base) ::simt::canary_set(base + off, 256, count * sizeof(T), file, line);
            off += 256;
        }
        return p;"""
TAKE_OLD = "T *take(size_t count, size_t align = 256) {"
TAKE_NEW = ("T *take(size_t count, size_t align = 256, const char *file = __builtin_FILE(), "
            "int line = __builtin_LINE()) {")


CARVE_FILL = re.compile(r"launch_fill_u32\((\w+), 0, cv\.off / 4, st\)")
SCAN_WS_OLD = "    return tiles * 8 + 512;\n}"
SCAN_WS_NEW = "    return tiles * 8 + 512 + (::simt::canary_on() ? 512 : 0);\n}"


def rewrite(text: str) -> str:
    text = text.replace('#include "../../include/', '#include "')
    # The GPU build keeps the serving loop switched off; the CPU build runs it.
    text = re.sub(r'\n\s*if \(cfg\) return set_error\(QO_ERR_NOT_FOUND, "the serving loop is disabled[^\n]*', "", text)
    # Carver: while canaries are enabled, 256 bytes of pattern after every
    # buffer (in the sizing pass too, so sizes and offsets stay consistent).
    text = text.replace(CARVE_OLD, CARVE_NEW).replace(TAKE_OLD, TAKE_NEW)
    # A fill over a whole carve (several buffers zeroed at once) retires the
    # canaries between them first, and the scan workspace leaves room for its
    # two canaries.
    text = CARVE_FILL.sub(r"launch_fill_u32(\1, 0, (::simt::canary_retire(\1, cv.off), cv.off / 4), st)", text)
    text = text.replace(SCAN_WS_OLD, SCAN_WS_NEW)
    text = EXTERN_LDS.sub(lambda m: f"{m.group(1)} *{m.group(2)} = reinterpret_cast<{m.group(1)} *>(::simt::dyn_lds());",
                          text)
    return rewrite_launches(text)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(" ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)


def _stale(target: Path, deps) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build(drivers=()) -> Path:
    SRC.mkdir(parents=True, exist_ok=True)
    OUT.mkdir(parents=True, exist_ok=True)
    inputs = [f for f in CSRC.iterdir() if f.suffix in (".hip", ".hpp")]
    inputs += [HERE / "simt.cpp", HERE / "hip" / "hip_runtime.h", Path(__file__), ROOT / "include" / "qo.h"]
    lib = OUT / "libqo_cpu.a"
    if not _stale(lib, inputs):
        for d in drivers:
            exe = OUT / d
            if _stale(exe, [lib, HERE / f"{d}.cpp"]):
                _run([CXX, *FLAGS, str(HERE / f"{d}.cpp"), str(lib), "-lpthread", "-o", str(exe)])
        return lib
    units = []
    for f in sorted(CSRC.iterdir()):
        if f.suffix not in (".hip", ".hpp"):
            continue
        dst = SRC / (f.stem + (".cpp" if f.suffix == ".hip" else ".hpp"))
        new = rewrite(f.read_text())
        if not dst.exists() or dst.read_text() != new:
            dst.write_text(new)
        if f.suffix == ".hip":
            units.append(dst)
    units.append(HERE / "simt.cpp")

    def comp(u: Path) -> Path:
        o = OUT / (u.stem + ".o")
        _run([CXX, *FLAGS, "-c", str(u), "-o", str(o)])
        return o

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        objs = list(ex.map(comp, units))
    if lib.exists():
        lib.unlink()
    _run(["ar", "rcs", str(lib), *map(str, objs)])
    for d in drivers:
        _run([CXX, *FLAGS, str(HERE / f"{d}.cpp"), str(lib), "-lpthread", "-o", str(OUT / d)])
    if VARIANT == "fast":
        # A shared library with libqo's C ABI, for host programs that load it by path.
        _run([CXX, *FLAGS, "-shared", "-o", str(OUT / "libqo_cpu.so"), *map(str, objs), "-lpthread"])
    return lib


if __name__ == "__main__":
    build(sys.argv[1:])
    print(f"built {OUT}")
