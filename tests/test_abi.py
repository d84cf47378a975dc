"""The C-ABI library loads and exports every function include/*.h declares.

CPU-only: no compute call is made (there is no GPU in the build container).
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#[^\n]*", "", text)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            prefix = text[max(0, m.start() - 80):m.start()]
            if name in ("sizeof", "if", "return"):
                continue
            if re.search(r"typedef[^;]*$", prefix):
                continue
            names.add(name)
    return sorted(names)


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ("cublas_2nn_sift", "cublas_2nn_sift_batch", "device_brute", "cublas_2nn_f",
                "cublas_2nn_brute_f", "lsh_gpu"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    import matching_sift_amd
    L = matching_sift_amd.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, "not exported: %s" % missing


def test_symbols_have_c_linkage():
    from matching_sift_amd import LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in declared_functions():
        assert n in exported, n  # unmangled


def test_python_signatures_cover_header():
    from matching_sift_amd._lib import SIGNATURES
    missing = [n for n in declared_functions() if n not in SIGNATURES]
    assert not missing, missing


def test_product_does_not_link_the_oracle():
    from matching_sift_amd import LIB_PATH
    out = subprocess.run(["readelf", "-d", LIB_PATH], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", LIB_PATH], stdout=subprocess.PIPE, text=True,
                          check=True).stdout
    assert "oracle_" not in syms


def test_gpu_code_object_is_gfx950():
    from matching_sift_amd import LIB_PATH
    data = open(LIB_PATH, "rb").read()
    assert b"gfx950" in data
