"""gravitas_amd -- MI355X (gfx950) Barnes-Hut force step for the gravitas engine.

The product is libgravitas_hip.so (C ABI in include/gravitas_hip.h); this
package builds it (``gravitas_amd.build``) and mirrors the reference's
``bh.Solver`` interface on top of it (``gravitas_amd.solver``), and the SPH
hydro's ``sph.Solver`` (``gravitas_amd.hydro``).
"""
from .hydro import GasParams, GasState, Hydro, HydroOut  # noqa: F401
from .solver import DeviceState, Params, Retard, Solver, accelerations_noalloc, gpu_available, gpu_name  # noqa: F401
