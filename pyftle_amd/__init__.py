"""pyftle_amd: MI355X-native FTLE engine behind pyFTLE's strategy API.

Drop-in names (pyFTLE 1.0.3 ``src/pyftle``):
``NeighboringParticles``, ``create_interpolator`` / ``HighPerfInterpolator`` /
``AnalyticalInterpolator``, ``create_integrator`` and the Euler/AB2/RK4
integrators, ``compute_flow_map_jacobian_2x2/3x3``, ``compute_cauchy_green_*``,
``max_eigenvalue_*``, ``compute_ftle_*``, ``FTLESolver``, the ``BatchSource``
implementations, ``AnalyticalSolver``, the ``config.yaml`` keys (``load_config``,
``FileFTLEManager``) and the ``.mat`` writer.  Additions: ``BatchPlan`` (many FTLE
windows in one persistent launch) and the learned flow-map surrogate
(``FlowMapMLP``, ``FlowMapEnsemble``, ``SurrogateFTLE``).  Compute runs in hand-written HIP
kernels (``libpyftle_hip.so``); there is no CPU fallback.
"""

from .particles import NeighboringParticles  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # Lazy imports keep `import pyftle_amd` cheap and torch-free until needed.
    import importlib

    table = {
        "AnalyticalSolver": ".api",
        "FTLESolver": ".ftle_solver",
        "create_integrator": ".integrate",
        "EulerIntegrator": ".integrate",
        "AdamsBashforth2Integrator": ".integrate",
        "RungeKutta4Integrator": ".integrate",
        "create_interpolator": ".interpolate",
        "HighPerfInterpolator": ".interpolate",
        "AnalyticalInterpolator": ".interpolate",
        "compute_flow_map_jacobian_2x2": ".cauchy_green",
        "compute_flow_map_jacobian_3x3": ".cauchy_green",
        "compute_cauchy_green_2x2": ".ftle",
        "compute_cauchy_green_3x3": ".ftle",
        "max_eigenvalue_2x2": ".ftle",
        "max_eigenvalue_3x3": ".ftle",
        "compute_ftle_2x2": ".ftle",
        "compute_ftle_3x3": ".ftle",
        "FileBatchSource": ".data_source",
        "AnalyticalBatchSource": ".data_source",
        "ArrayBatchSource": ".data_source",
        "FileFTLEManager": ".experiment",
        "load_config": ".experiment",
        "create_writer": ".file_writers",
        "BatchPlan": ".batch",
        "Window": ".batch",
        "FlowMapMLP": ".surrogate",
        "FlowMapEnsemble": ".surrogate",
        "SurrogateFTLE": ".surrogate_ftle",
    }
    if name in table:
        mod = importlib.import_module(table[name], __name__)
        return getattr(mod, name)
    raise AttributeError(name)
