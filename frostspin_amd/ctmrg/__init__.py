from .sequential_ctmrg import SequentialCTMRG

__all__ = ["SequentialCTMRG"]
