"""sparrowpy_amd: sparrowpy's directional radiosity hot path on AMD MI355X.

``DirectionalRadiosityFast`` keeps the reference class's API
(``bake_geometry``, ``init_source_energy``, ``calculate_energy_exchange``,
``collect_energy_receiver_mono``/``_patchwise``) with every numerical step in
hand-written HIP kernels (``libsparrow_hip.so``, C ABI in
``include/sparrow_hip.h``).  Host-side input producers (wall subdivision,
BRDF construction, direction samplings) mirror sparrowpy's modules of the
same names.
"""
from . import brdf, geometry, samplings, scenes, testing
from .coordinates import Coordinates, FrequencyData, TimeData
from .geometry import Polygon

__version__ = "0.1.0"


def __getattr__(name):
    # the GPU class pulls in torch; import it on first use
    if name == "DirectionalRadiosityFast":
        from .radiosity import DirectionalRadiosityFast
        return DirectionalRadiosityFast
    if name in ("ops", "radiosity"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)


__all__ = ["DirectionalRadiosityFast", "Coordinates", "FrequencyData",
           "TimeData", "Polygon", "brdf", "geometry", "samplings", "scenes",
           "testing"]
