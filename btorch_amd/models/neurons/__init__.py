from .glif import GLIF3
from .lif import IF, LIF

__all__ = ["LIF", "IF", "GLIF3"]
