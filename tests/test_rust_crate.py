"""The Rust crate under rmetis_amd/rust (the reference's own host language)
cannot be compiled in this image (no cargo or rustc), so its binding layer is
checked here as text: every extern "C" function it declares is declared in
include/rmetis.h with the same number of parameters and is exported by the
built library, and build.rs compiles the same sources as rmetis_amd/build.py."""
import ctypes
import os
import re

import pytest

from rmetis_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "rmetis_amd", "rust")


def _read(*parts):
    with open(os.path.join(*parts)) as fh:
        return fh.read()


def _params(body):
    body = body.strip()
    if not body or body == "void":
        return 0
    depth, n = 0, 1
    for ch in body:
        depth += ch in "(<[" and 1 or 0
        depth -= ch in ")>]" and 1 or 0
        n += ch == "," and depth == 0
    return n - body.rstrip().endswith(",")


def _rust_decls():
    src = _read(CRATE, "src", "ffi.rs")
    block = src[src.index('extern "C" {'):]
    return {m.group(1): _params(m.group(2))
            for m in re.finditer(r"pub fn (\w+)\(([^)]*)\)", block, re.S)}


def _c_decls():
    src = _read(ROOT, "include", "rmetis.h")
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return {m.group(1): _params(m.group(2))
            for m in re.finditer(r"\b((?:RMETIS_|rmetis_)\w+)\(([^;{]*?)\)\s*;", src, re.S)}


def test_every_rust_binding_matches_the_header():
    rust, c = _rust_decls(), _c_decls()
    assert {"RMETIS_PartGraphKway", "RMETIS_PartGraphRecursive", "RMETIS_NodeND"} <= set(rust)
    for name, n in rust.items():
        assert name in c, name
        assert n == c[name], (name, n, c[name])


def test_build_rs_compiles_the_library_sources():
    src = _read(CRATE, "build.rs")
    block = src[src.index("const SOURCES"):src.index("];")]
    assert re.findall(r'"([\w.]+)"', block) == B.SOURCES
    assert "gfx950" in src and "-packed-fp32-ops" in src


def test_rust_bindings_are_exported():
    if not os.path.exists(B.LIB):
        pytest.skip("librmetis_hip.so is not built")
    lib = ctypes.CDLL(B.LIB)
    for name in _rust_decls():
        assert hasattr(lib, name), name
