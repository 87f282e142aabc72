from .glif import GLIF3
from .lif import IF, LIF
from .mixed import MixedNeuronPopulation
from .two_compartment import TwoCompartmentGLIF

__all__ = ["LIF", "IF", "GLIF3", "TwoCompartmentGLIF", "MixedNeuronPopulation"]
