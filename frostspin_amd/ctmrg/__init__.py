from .sequential_ctmrg import SequentialCTMRG
from .simultaneous_ctmrg import SimultaneousCTMRG
from .symmetrized_ctmrg import SymmetrizedCTMRG

__all__ = ["SequentialCTMRG", "SimultaneousCTMRG", "SymmetrizedCTMRG"]
