"""Replay of an explicit compiled tree-transform structure (oracle).

Restates tenet-operations/src/transform_replay.rs:3413-3606 (Single strided
axpby, Multi pack -> src * U^T -> scatter, :4428-4822) and the inactive
destination scaling of :288-330 over caller-given block layouts, so the
reference's integer known-answer tests (tenet-tensors/src/tests/mod.rs:1099-1320,
tree_transform_exec.rs:184-600) can be replayed exactly.

A layout is (offset, shape, strides) with column-major element order; a
Multi group's blocks must share element_count (shapes may differ: elements
correspond by column-major linear index in the destination axis order, as in
the reference's packed column).  TEST INFRASTRUCTURE ONLY.
"""
import numpy as np


def packed_column_major(shapes):
    """BlockStructure::packed_column_major: consecutive dense blocks."""
    out, off = [], 0
    for shp in shapes:
        strides, st = [], 1
        for d in shp:
            strides.append(st)
            st *= d
        out.append((off, tuple(shp), tuple(strides)))
        off += st
    return out


def _gather(data, layout):
    off, shape, strides = layout
    n = int(np.prod(shape)) if shape else 1
    idx = np.full(n, off, dtype=np.int64)
    # column-major enumeration: first axis fastest
    lin = np.arange(n)
    for d, s in zip(shape, strides):
        idx += (lin % d) * s
        lin //= d
    return idx


def replay(dst, src, dst_layouts, src_layouts, specs, alpha, beta, overwrite=False):
    """dst <- alpha * T(src) + beta * dst (or overwrite) for explicit specs.

    specs: ('single', d, s, coeff) or ('multi', [d...], [s...], U_dst_src flat row-major).
    src layouts for a Single/Multi are read in the DESTINATION axis order:
    the caller passes src strides already permuted.
    """
    dst = np.array(dst, dtype=np.float64)
    src = np.asarray(src, dtype=np.float64)
    touched = set()
    updates = {}
    for spec in specs:
        if spec[0] == "single":
            _, d, s, c = spec
            updates[d] = c * src[_gather(src, src_layouts[s])]
            touched.add(d)
        else:
            _, ds, ss, U = spec
            U = np.asarray(U, dtype=np.float64).reshape(len(ds), len(ss))
            P = np.stack([src[_gather(src, src_layouts[s])] for s in ss], axis=1)
            Q = P @ U.T
            for j, d in enumerate(ds):
                updates[d] = Q[:, j]
                touched.add(d)
    for d, lay in enumerate(dst_layouts):
        idx = _gather(dst, lay)
        if d in touched:
            base = 0.0 if overwrite else beta * dst[idx]
            dst[idx] = alpha * updates[d] + base
        else:
            dst[idx] = 0.0 if overwrite else beta * dst[idx]
    return dst
