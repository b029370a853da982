from .simple_update import SimpleUpdate

__all__ = ["SimpleUpdate"]
