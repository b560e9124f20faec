"""The C ABI's struct layouts and enum values against the Rust binding.

The repository's integration guide holds, in a `hip_ffi.rs` code block, the
`#[repr(C)]` structs and constants a TeNeT maintainer would add
(tenet-dense/src/hip_ffi.rs).  This test parses that block, computes each
struct's C layout from the Rust field types (size,
alignment, field offsets under the C rules `#[repr(C)]` follows), and
compares it with what gcc reports for include/tenet_hip.h.  The ctypes
mirror in tenet_amd/_native.py is checked against the same numbers, and the
tnt_op values against the reference's MatrixOp declaration order
(tenet-dense/src/executor.rs:24-31: Identity, Transpose, Adjoint).
No GPU is needed.
"""
import ctypes as C
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from tenet_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# size and alignment of the Rust field types used by the binding (x86-64)
RUST_TYPES = {"usize": (8, 8), "isize": (8, 8), "i64": (8, 8), "u64": (8, 8), "f64": (8, 8),
              "i32": (4, 4), "u32": (4, 4), "c_int": (4, 4)}


def rust_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```rust\n// tenet-dense/src/hip_ffi\.rs.*?```", text, re.S)
    assert m, "INTEGRATION.md has no hip_ffi.rs block"
    return m.group(0)


def rust_structs():
    """name -> [(field, size, align)] of every #[repr(C)] struct with fields."""
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\](?:\s*#\[derive\([^)]*\)\])?\s*pub struct (\w+)\s*\{(.*?)\}", rust_block(), re.S):
        name, body = m.group(1), m.group(2)
        if "_p:" in body:
            continue  # opaque handle
        fields = []
        for f in re.finditer(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", body.strip()):
            fname, ty = f.group(1), f.group(2).strip()
            if ty.startswith("*"):
                size, align = 8, 8
            elif ty.startswith("["):
                a = re.match(r"\[(\w+);\s*(\d+)\]", ty)
                es, ea = RUST_TYPES[a.group(1)]
                size, align = es * int(a.group(2)), ea
            else:
                size, align = RUST_TYPES[ty]
            fields.append((fname, size, align))
        out[name] = fields
    return out


def c_layout(fields):
    """(sizeof, {field: offset}) under the C / #[repr(C)] rules."""
    off, align_max, offs = 0, 1, {}
    for name, size, align in fields:
        off = (off + align - 1) // align * align
        offs[name] = off
        off += size
        align_max = max(align_max, align)
    return (off + align_max - 1) // align_max * align_max, offs


def rust_consts():
    return {m.group(1): int(m.group(2)) for m in
            re.finditer(r"pub const (TNT_\w+): \w+ = (-?\d+);", rust_block())}


@pytest.fixture(scope="module")
def gcc_layout():
    """sizeof / offsetof / enum values from the C header, printed by gcc."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs, consts = rust_structs(), rust_consts()
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tenet_hip.h"', "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'  printf("S {name} %zu\\n", sizeof({name}));')
        for f, _, _ in fields:
            lines.append(f'  printf("F {name} {f} %zu\\n", offsetof({name}, {f}));')
    for c in consts:
        lines.append(f'  printf("C {c} %d\\n", (int)({c}));')
    lines.append("  return 0;\n}")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "abi.c"), os.path.join(d, "abi")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    sizes, offsets, values = {}, {}, {}
    for ln in out.splitlines():
        p = ln.split()
        if p[0] == "S":
            sizes[p[1]] = int(p[2])
        elif p[0] == "F":
            offsets[(p[1], p[2])] = int(p[3])
        else:
            values[p[1]] = int(p[2])
    return sizes, offsets, values


def test_rust_block_covers_every_descriptor_struct():
    names = set(rust_structs())
    assert {"tnt_gemm_job", "tnt_gemm_ptr_job", "tnt_layout", "tnt_transform_block", "tnt_transform_desc",
            "tnt_region", "tnt_contract_desc"} <= names


def test_rust_repr_c_layout_matches_header(gcc_layout):
    sizes, offsets, _ = gcc_layout
    for name, fields in rust_structs().items():
        size, offs = c_layout(fields)
        assert sizes[name] == size, (name, sizes[name], size)
        for f, o in offs.items():
            assert offsets[(name, f)] == o, (name, f, offsets[(name, f)], o)


def test_rust_constants_match_header(gcc_layout):
    _, _, values = gcc_layout
    consts = rust_consts()
    assert consts, "no constants parsed"
    for c, v in consts.items():
        assert values[c] == v, (c, values[c], v)


def test_matrix_op_order_matches_reference():
    # MatrixOp { Identity, Transpose, Adjoint } in declaration order
    consts = rust_consts()
    assert (consts["TNT_OP_IDENTITY"], consts["TNT_OP_TRANSPOSE"], consts["TNT_OP_ADJOINT"]) == (0, 1, 2)
    assert (N.OP_IDENTITY, N.OP_TRANSPOSE, N.OP_ADJOINT) == (0, 1, 2)
    assert (N.F64, N.F32, N.C32, N.C64) == (consts["TNT_F64"], consts["TNT_F32"], consts["TNT_C32"],
                                            consts["TNT_C64"])


CTYPES = {"tnt_gemm_job": N.GemmJob, "tnt_gemm_ptr_job": N.GemmPtrJob, "tnt_layout": N.Layout,
          "tnt_transform_block": N.TransformBlock, "tnt_transform_desc": N.TransformDesc,
          "tnt_region": N.Region, "tnt_contract_desc": N.ContractDesc}


@pytest.mark.parametrize("name", sorted(CTYPES))
def test_ctypes_mirror_matches_header(gcc_layout, name):
    sizes, offsets, _ = gcc_layout
    cls = CTYPES[name]
    assert C.sizeof(cls) == sizes[name]
    for f, _ in cls._fields_:
        assert getattr(cls, f).offset == offsets[(name, f)], (name, f)
