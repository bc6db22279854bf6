"""specseqplusplus_amd — MI355X-native persistent-homology reduction engine (libssq).

The compute path is libssq.so (HIP kernels for gfx950 behind the C ABI in include/ssq.h);
this package is its Python host side.
"""
from .api import Handle, Matrix, collapse_edges, load_phat, params, reduce, save_pairs, vr_build  # noqa: F401
from .specseq import SpecSeqPlusPlus  # noqa: F401
from ._lib import LIB_PATH, SSQError  # noqa: F401
