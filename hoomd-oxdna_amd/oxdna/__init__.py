"""oxDNA1 / oxDNA2 / oxRNA2 on AMD Instinct MI355X (gfx950).

The reference's Python layer (``io``, ``forces``, ``model``) over a HIP force engine
reached through a C ABI (include/oxdna_hip.h, liboxdna_hip.so). ``sim`` provides the
small host-engine surface the reference's workflows use (snapshot -> simulation ->
integrator with forces -> run), and ``ensemble`` the batched replica runner.

Importing the package does not need a GPU; building a simulation does, and fails
loudly when the HIP library is missing.
"""

from . import io  # noqa: F401
from . import forces  # noqa: F401
from . import sim  # noqa: F401
from . import model  # noqa: F401
from . import systems  # noqa: F401
from . import ensemble  # noqa: F401

version = "0.1.0"

__all__ = ["io", "forces", "sim", "model", "systems", "ensemble", "version"]
