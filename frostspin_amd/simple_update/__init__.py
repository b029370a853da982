from .simple_update import SimpleUpdate
from .symmetrized_simple_update import SymmetrizedSimpleUpdate

__all__ = ["SimpleUpdate", "SymmetrizedSimpleUpdate"]
