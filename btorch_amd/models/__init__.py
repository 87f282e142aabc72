"""Module API mirroring btorch.models for the recurrent SNN time step."""

from . import environ, functional
from .base import BaseNode, MemoryModule, ParamBufferMixin, StepModule
from .history import SpikeHistory
from .linear import BaseSparseConn, SparseConn, available_sparse_backends
from .neurons import GLIF3, IF, LIF
from .rnn import RecurrentNN, RecurrentNNAbstract, make_rnn
from .surrogate import Erf, Sigmoid, SuperSpike, SurrogateFunctionBase
from .synapse import BasePSC, DelayedPSC, ExponentialPSC, HeterSynapsePSC

__all__ = [
    "environ", "functional", "BaseNode", "MemoryModule", "ParamBufferMixin", "StepModule",
    "SpikeHistory", "BaseSparseConn", "SparseConn", "available_sparse_backends", "IF", "LIF", "GLIF3",
    "RecurrentNN", "RecurrentNNAbstract", "make_rnn", "Erf", "Sigmoid", "SuperSpike",
    "SurrogateFunctionBase", "BasePSC", "DelayedPSC", "ExponentialPSC", "HeterSynapsePSC",
]
