"""Host-side mirror of the reference engine class SpecSeqPlusPlus (specseq.h:63-1194).

Same constructor arguments and call sequence: construct from per-column boundaries and
dimensions, GPU_Init(), GPU_Compute(), GetPersistencePairs().  The work is done by libssq in the
reference's two phases: GPU_Init (ssq_init: upload, edge collapse, high-dim clearing or the
anti-transpose, init passes, pending lists) and GPU_Compute (ssq_compute: the page loop and the
pairs).  Each returns its phase's seconds, as the reference's do (main.cu:337-348).
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
        self._essential = None
        self._stats = None
        self._initialized = False
        self.DataFileName = ""

    def setDataFileName(self, name):
        self.DataFileName = name

    def GPU_Init(self):
        """specseq.h:513-590; seconds of the phase."""
        self._stats = self._handle.init(self.matrix, self.params)
        self._initialized = True
        return self._stats["t_init_ms"] / 1000.0

    def GPU_Compute(self):
        """specseq.h:592-1060; seconds of the phase.  Runs GPU_Init first if it has not run."""
        if not self._initialized:
            self.GPU_Init()
        pairs, ess, st = self._handle.compute(stats=True)
        self._initialized = False
        self._pairs, self._essential, self._stats = pairs, ess, st
        return st["t_reduce_ms"] / 1000.0

    def GetPersistencePairs(self):
        """specseq.h:1182-1193: (birth, death) pairs in birth order."""
        if self._pairs is None:
            self.GPU_Compute()
        return [(int(b), int(d)) for b, d in self._pairs]

    def GetEssential(self):
        """Cells that stay unpaired (the output's essential classes)."""
        if self._pairs is None:
            self.GPU_Compute()
        return [int(c) for c in self._essential]
