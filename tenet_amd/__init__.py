"""tenet_amd: MI355X (gfx950) backend for TeNeT's block-sparse symmetric contraction path.

Native pieces (built in-tree by ``tenet_amd.build``):
  lib/libtenet_hip.so   recoupling replay + grouped MFMA block GEMM + reductions (C ABI,
                        include/tenet_hip.h)
  lib/libtenet_host.so  host structure compiler (include/tenet_host.h)
"""
from .tensor import GradedSpace, HomSpace, Rule, Runtime, TensorMap, TenetError  # noqa: F401
