"""btorch_amd: MI355X-native recurrent spiking-network time step.

A drop-in for btorch's RSNN hot path -- ``btorch_amd.models`` mirrors
``btorch.models`` (neurons, synapses, SpikeHistory, SparseConn, RecurrentNN,
functional, environ) and ``btorch_amd.config.sparse`` mirrors
``btorch.config.sparse`` -- with every step computed by hand-written gfx950
HIP kernels in ``libbtx.so`` (C ABI: include/btorch_amd.h).
"""

__version__ = "0.1.0"

from . import config, connectome, models  # noqa: E402

__all__ = ["config", "connectome", "models", "__version__"]
