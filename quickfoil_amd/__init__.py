"""quickfoil_amd: QuickFOIL's candidate-literal evaluation path on MI355X (gfx950).

The product is the C-ABI library libquickfoil_amd.so (include/quickfoil_amd.h) built from the
HIP sources in csrc/; this package exposes it to Python through ctypes.  There is no CPU
fallback: calls raise if the library is not built or no GPU is present.
"""
from ._native import (Context, LiteralArray, TableSet, QfError, load_library, LIB_PATH, MODES, MODE_BATCHED,
                      MODE_PLAIN_GROUP, MODE_PLAIN_LITERAL, SIGNATURES)

__all__ = ["Context", "LiteralArray", "TableSet", "QfError", "load_library", "LIB_PATH", "MODES", "MODE_BATCHED",
           "MODE_PLAIN_GROUP", "MODE_PLAIN_LITERAL", "SIGNATURES"]
