"""Host-side mirror of the reference engine class SpecSeqPlusPlus (specseq.h:63-1194).

Same constructor arguments and call sequence: construct from per-column boundaries and
dimensions, GPU_Init(), GPU_Compute(), GetPersistencePairs().  The work is done by libssq;
GPU_Init uploads and runs the pre-passes, GPU_Compute runs the page loop (both happen inside
one ssq_reduce call, so GPU_Init only validates and GPU_Compute returns the call's time).
"""
import numpy as np

from .api import Handle, Matrix, params


class SpecSeqPlusPlus:
    def __init__(self, hv_csc, hv_dim, level_capacity, dualize=0, model=0, highclear_maxneighbors=10,
                 ss_block_size=102400, block_size_growth_rate=1.05, queue_size=65536, device=0):
        """hv_csc: list of boundary index lists (0-based), hv_dim: cell dimensions,
        level_capacity: (level1, level2) column-store slot capacities (specseq.h:486)."""
        cols = [sorted(set(int(x) for x in c)) for c in hv_csc]
        colptr = np.zeros(len(cols) + 1, dtype=np.int64)
        colptr[1:] = np.cumsum([len(c) for c in cols]) if cols else []
        rows = np.array([x for c in cols for x in c], dtype=np.int32)
        self.matrix = Matrix(colptr, rows, np.asarray(hv_dim, dtype=np.int32))
        lc = list(level_capacity)
        if len(lc) == 4:  # the reference passes int32_t level_capacity[4] with [1], [2] used
            lc = lc[1:3]
        self.params = params(dualize=int(dualize), model=int(model), max_iteration=int(highclear_maxneighbors),
                             ss_block_size=int(ss_block_size), growth=float(block_size_growth_rate),
                             queue_size=int(queue_size), level1=int(lc[0]), level2=int(lc[1]))
        self._handle = Handle(device)
        self._pairs = None
        self._stats = None
        self.DataFileName = ""

    def setDataFileName(self, name):
        self.DataFileName = name

    def GPU_Init(self):
        return 0.0

    def GPU_Compute(self):
        pairs, _ess, st = self._handle.reduce(self.matrix, self.params, stats=True)
        self._pairs = pairs
        self._stats = st
        return st["t_total_ms"] / 1000.0

    def GetPersistencePairs(self):
        if self._pairs is None:
            self.GPU_Compute()
        return [(int(b), int(d)) for b, d in self._pairs]
