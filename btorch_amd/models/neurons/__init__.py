from .lif import IF, LIF

__all__ = ["LIF", "IF"]
