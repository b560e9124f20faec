"""CPU oracle for TeNeT's block-sparse symmetric contraction path.

TEST INFRASTRUCTURE ONLY.  This package restates, in plain Python + numpy and
in FP64, the reference's host algorithm for the path (sector providers,
fusion-tree enumeration, coupled-sector block layout, fusion-tree-pair
recoupling coefficients, tree-transform group planning, per-sector block GEMM,
weighted reductions).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline; the product path (``tenet_amd``) never calls it.

Every function cites the reference file:line it follows (paths relative to
the reference repository root).

Pinning (what fixes each part independently of this restatement):

* layout / fill        -- microbench checksums (docs/su2_authority.md:81-83),
                          TensorKit repartition fixtures and invariant stream
                          (benchmarks/tensorkit_semantic_oracle.out),
                          closed-form offsets (tests/test_oracle_stream.py).
* SU(2) F/R/dim/FS     -- the 181 TensorKit F values and R/dim/FS tables,
                          pentagon and hexagon identities, F orthogonality.
* braid / bend coeffs  -- SU(2) KAT 0.5 / 0.8660254037844386 / 53 / 92.6025...
                          (tenet-tensors/src/tests/contract_fusion.rs:4693-4776),
                          TensorKit invariant stream s3/s4 and repartition data.
* replay (Single/Multi)-- integer KATs (tenet-tensors/src/tests/mod.rs:1099-1320,
                          tree_transform_exec.rs:184-600).
* contraction          -- invariant stream s2/s4/s5/s7/s8, quickstart 50.0,
                          brute-force dense contraction on tiny U(1) inputs.
"""
