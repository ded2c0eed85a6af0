"""rendiation_amd -- MI355X-native ray-query hot path of rendiation.

Host side of librdrt.so: BVH over triangle soups with closest-hit ray
queries on gfx950.  The compute path is the HIP library; this package holds
its Python bindings (``acceleration``), and seeded synthetic inputs
(``synth``).
"""

from .acceleration import (  # noqa: F401
    AccelerationStructureSystem,
    BlasHandle,
    BuildOptions,
    TlasHandle,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
    trace_closest,
)
