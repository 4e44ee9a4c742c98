"""CPU checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host-only planning matches the reference's tree
schedule (alg/QR/tsqr/bitree_tsqr.cxx:172-208).  No compute call is made."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from candmc_amd import _lib
from oracle import candmc_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "candmc_qr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cqr_[A-Za-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES)


def test_status_strings():
    lib = _lib.load()
    assert lib.cqr_status_string(0) == b"ok"
    assert lib.cqr_status_string(2) == b"unsupported shape"


def _plan(m, b, hint=0, rows=0):
    lib = _lib.load()
    L, R, V = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    s = lib.cqr_plan_tsqr(m, b, hint, rows, ctypes.byref(L), ctypes.byref(R), ctypes.byref(V))
    return s, L.value, R.value, V.value


@pytest.mark.parametrize("m,b,hint", [(8192, 64, 4), (32768, 128, 0), (136, 17, 4), (1000, 50, 3),
                                      (16384, 16, 0), (300, 128, 0), (128, 128, 0)])
def test_plan_matches_reference_tree(m, b, hint):
    s, L, rows, levels = _plan(m, b, hint)
    assert s == 0
    assert L >= max(hint, 1) and L <= m // b
    assert rows == -(-m // L)
    assert levels == len(O.bitree_levels(L))
    # every leaf keeps at least b rows
    assert m // L >= b


def test_plan_rejects_bad_shapes():
    assert _plan(10, 20)[0] == 1        # m < b
    assert _plan(1000, 129)[0] == 2     # b > 128


def test_tree_data_size():
    lib = _lib.load()
    s, L, _, _ = _plan(8192, 64, 4)
    nodes = sum(len(O.bitree_pairs(n)) for n in O.bitree_levels(L))
    assert lib.cqr_tree_data_doubles(8192, 64, 4) == L * 64 * 64 + nodes * (2 * 64 * 64 + 64)


def test_compute_entry_points_reject_null_context():
    lib = _lib.load()
    assert lib.cqr_hh_recon_qr(None, None, 1, 1, 1, None, 0) == 1
    assert lib.cqr_qr_2d(None, None, 1, 1, 1, 1, None, 0) == 1
    assert lib.cqr_apply_tsqr_QT(None, None, 1, None, None, 1, 1, 1, 1, 0, None) == 1
    assert lib.cqr_apply_tsqr_Q(None, None, 1, None, None, 1, 1, 1, 1, 0, None) == 1
    assert lib.cqr_construct_Q1_id(None, None, 1, None, None, 1, 1, 1, 1, 0, None, None, 0) == 1
    assert lib.cqr_qr_tree_2d(None, None, 1, 1, 1, 1, -1, 0) == 1
    assert lib.cqr_tree_data_export(None, None, 1, 1, 0, 0, None, None) == 1


def _c_kind(decl):
    """'ptr' for any pointer, else the base C type of one declaration."""
    if "*" in decl:
        return "ptr"
    words = [w for w in decl.split() if w not in ("const", "unsigned")]
    return words[0] if len(words) <= 2 else "?"


def _ctypes_kind(t):
    if t in (ctypes.c_void_p, ctypes.c_char_p) or (isinstance(t, type) and issubclass(t, ctypes._Pointer)):
        return "ptr"
    return {ctypes.c_int64: "int64_t", ctypes.c_int: "int", ctypes.c_size_t: "size_t",
            ctypes.c_double: "double", ctypes.c_char: "char"}.get(t, repr(t))


def test_ctypes_signatures_match_header_prototypes():
    """Every prototype's argument count, argument kinds and return kind
    agree with the ctypes table, so no call passes a 32-bit int where the
    library reads 64 bits."""
    src = open(os.path.join(ROOT, "include", "candmc_qr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    protos = re.findall(r"([A-Za-z_][A-Za-z0-9_ \*]*?)\b(cqr_[A-Za-z0-9_]+)\s*\(([^;{]*?)\)\s*;", src)
    assert {name for _, name, _ in protos} == set(header_symbols())
    for ret, name, args in protos:
        decls = [a.strip() for a in args.split(",") if a.strip() and a.strip() != "void"]
        res, argtypes = _lib.SIGNATURES[name]
        assert [_c_kind(d) for d in decls] == [_ctypes_kind(t) for t in argtypes], name
        assert _c_kind(ret.strip()) == _ctypes_kind(res), name


def test_plan_sweep_against_reference_tree():
    """Seeded sweep over shapes, width and leaf hints: every accepted plan
    keeps ≥ b rows per leaf, honours the hint where the rows allow it, and
    its level count and tree_data size follow the reference's tree
    (bitree_tsqr.cxx:172-208)."""
    lib = _lib.load()
    rng = np.random.default_rng(1234)
    accepted = 0
    for _ in range(300):
        b = int(rng.integers(1, 129))
        m = int(rng.integers(b, 200 * b + 1))
        hint = int(rng.integers(0, 40))
        s, L, rows, levels = _plan(m, b, hint)
        assert s in (0, 2)
        if s:
            continue
        accepted += 1
        assert 1 <= L <= max(1, m // b)
        assert L >= min(max(hint, 1), m // b)
        assert rows == -(-m // L) and m // L >= b
        assert levels == len(O.bitree_levels(L))
        nodes = sum(len(O.bitree_pairs(n)) for n in O.bitree_levels(L))
        assert lib.cqr_tree_data_doubles(m, b, hint) == L * b * b + nodes * (2 * b * b + b)
    assert accepted > 250


@pytest.mark.parametrize("L", list(range(1, 70)))
def test_reference_tree_absorbs_every_leaf_once(L):
    """The oracle's tree schedule: each level pairs receiver r with sender
    r + ceil(n/2); over all levels every leaf but 0 is sent exactly once,
    always to a lower rank still active, and leaf 0 ends holding R."""
    sent = []
    active = set(range(L))
    for n in O.bitree_levels(L):
        assert active == set(range(n))
        for r, src in O.bitree_pairs(n):
            assert r < src < n and r in active and src in active
            sent.append(src)
        active -= {src for _, src in O.bitree_pairs(n)}
    assert sorted(sent) == list(range(1, L))
    assert active == {0}


C_CLIENT = r"""
#include <stdio.h>
#include "candmc_qr.h"
int main(void) {
  int64_t L = 0, rows = 0, levels = 0;
  if (cqr_plan_tsqr(8192, 64, 4, 0, &L, &rows, &levels) != CQR_OK) return 1;
  printf("%lld %lld %lld %s\n", (long long)L, (long long)rows, (long long)levels, cqr_status_string(CQR_OK));
  return cqr_hh_recon_qr(NULL, NULL, 1, 1, 1, NULL, 0) == CQR_ERR_INVALID ? 0 : 2;
}
"""


def test_headers_compile_standalone_and_link_from_c(tmp_path):
    """candmc_qr.h is plain C99 (no HIP or torch types) and a C program
    links the library and calls its host-only entry points; the drop-in
    header is plain C++11."""
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("needs gcc and g++")
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "client.c"
    src.write_text(C_CLIENT)
    exe = str(tmp_path / "client")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I" + inc, str(src), "-o", exe,
                    "-L" + libdir, "-lcandmc_qr", "-Wl,-rpath," + libdir], check=True, timeout=120)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stdout + run.stderr
    s, L, rows, levels = _plan(8192, 64, 4)
    assert run.stdout.split() == [str(L), str(rows), str(levels), "ok"]
    cxx = tmp_path / "dropin.cpp"
    cxx.write_text('#include "candmc/CANDMC_QR.h"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I" + inc, str(cxx)],
                   check=True, timeout=120)
