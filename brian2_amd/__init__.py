"""brian2_amd: an MI355X (gfx950) `hip` target for Brian2's clock-driven step.

Hot path: Brian's stateupdate / threshold / reset code objects fused into one
generated kernel per group and step, wavefront-ballot spike compaction into a
device ring buffer indexed by integer delay, and a CSR-by-(source, delay)
gather with a deterministic per-target fold of the delivered events (no
floating-point atomics), replayed step after step from hipGraphs.

The C ABI (include/hipbrian.h) lives in libhipbrian.so next to this file; the
package refuses to import without it.
"""
from . import _lib  # noqa: F401  (fails loudly when the library is missing)
from .runtime import (  # noqa: F401
    Device,
    Model,
    Network,
    NeuronGroup,
    PoissonGroup,
    PopulationRateMonitor,
    SpikeGeneratorGroup,
    SpikeMonitor,
    StateMonitor,
    Subgroup,
    Synapses,
)

__version__ = "0.1"
