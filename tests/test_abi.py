"""CPU: the C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls are made here)."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "query_on_opencl_fpga_amd" / "libqo.so"
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions(text: str) -> set[str]:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"#[^\n]*", " ", text)
    names = set()
    # top-level statements only: drop struct/enum bodies first
    flat = re.sub(r"\{[^{}]*\}", " ", text)
    for stmt in flat.split(";"):
        stmt = " ".join(stmt.split())
        if not stmt or stmt.startswith("typedef") or stmt.startswith("extern \"C\""):
            stmt = stmt.replace('extern "C"', "").strip()
            if stmt.startswith("typedef") or not stmt:
                continue
        m = re.match(r"^(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", stmt)
        if m:
            names.add(m.group(1))
    return names


def test_headers_declare_functions():
    names = set()
    for h in HEADERS:
        names |= declared_functions(h.read_text())
    assert len(HEADERS) >= 2
    assert {"qo_hash_build", "qo_hash_probe", "qo_filter_agg", "qo_groupby_agg", "qo_hj_filter_sort",
            "qo_batch_create", "buildHashTable", "probeHashTable", "CL_RangeSelection", "CL_GroupBy"} <= names


@pytest.mark.skipif(not LIB.exists(), reason="libqo.so is not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = []
    for h in HEADERS:
        for name in sorted(declared_functions(h.read_text())):
            if not hasattr(lib, name):
                missing.append(f"{h.name}:{name}")
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="libqo.so is not built")
def test_version_and_status_strings_need_no_gpu():
    lib = ctypes.CDLL(str(LIB))
    lib.qo_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.qo_version()
