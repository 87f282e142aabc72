"""Module API mirroring btorch.models for the recurrent SNN time step."""

from . import environ, functional
from .base import BaseNode, MemoryModule, ParamBufferMixin, StepModule
from .history import SpikeHistory
from .linear import BaseSparseConn, SparseConn, SparseConstrainedConn, available_sparse_backends
from .neurons import GLIF3, IF, LIF, MixedNeuronPopulation, TwoCompartmentGLIF
from .rnn import (ApicalRecurrentNN, RecurrentNN, RecurrentNNAbstract, SomaApicalRecurrentNN,
                  make_rnn)
from .surrogate import ATan, Erf, Sigmoid, SuperSpike, SurrogateFunctionBase
from .synapse import (AlphaPSC, AlphaPSCBilleh, BasePSC, DelayedPSC, DualExponentialPSC,
                      ExponentialPSC, HeterSynapsePSC)

__all__ = [
    "environ", "functional", "BaseNode", "MemoryModule", "ParamBufferMixin", "StepModule",
    "SpikeHistory", "BaseSparseConn", "SparseConn", "SparseConstrainedConn",
    "available_sparse_backends", "IF", "LIF", "GLIF3",
    "TwoCompartmentGLIF", "MixedNeuronPopulation", "RecurrentNN", "RecurrentNNAbstract",
    "ApicalRecurrentNN", "SomaApicalRecurrentNN", "make_rnn", "ATan", "Erf", "Sigmoid", "SuperSpike",
    "SurrogateFunctionBase", "BasePSC", "DelayedPSC", "ExponentialPSC", "HeterSynapsePSC",
    "AlphaPSC", "AlphaPSCBilleh", "DualExponentialPSC",
]
