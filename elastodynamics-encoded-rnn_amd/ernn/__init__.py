"""ERNN on MI355X: the Elastodynamics-encoded RNN forward recurrence and its
reverse-time adjoint as hand-written HIP kernels for gfx950, behind the
reference's nn.Module interface (guoziye97/Elastodynamics-encoded-RNN,
Parameter_inversion/Inversion_adam.py and Damage_inversion/Inversion_adam.py).

The compute path is libernn_hip.so (C ABI: include/ernn.h) loaded with ctypes;
there is no pure-PyTorch fallback.
"""
from .problem import (PARAM_NAMES, XZ_NAMES, Shot, Survey, bond_rotate, cfl_dt, ellipse_damage,
                      gaussian_sources, max_phase_velocity, point_source, ricker, row_receivers)
from .engine import Engine
from .function import ERNNFunction, ernn_traces
from .simulator import (ERNN, CFRPWaveSimulatorRicker, SpatiallyVaryingAnisotropicSimulator,
                        total_variation_2d)
from .inversion import (CFRPParameterInverterRicker, SpatiallyVaryingParameterInverter, TraceInversion,
                        add_noise_rms, add_noise_snr)
from .optim import FusedTVAdam
from . import autotune, configs, data, fem

__all__ = [
    "PARAM_NAMES", "XZ_NAMES", "Shot", "Survey", "bond_rotate", "cfl_dt", "ellipse_damage",
    "gaussian_sources", "max_phase_velocity", "point_source", "ricker", "row_receivers",
    "Engine", "ERNNFunction", "ernn_traces", "ERNN", "CFRPWaveSimulatorRicker",
    "SpatiallyVaryingAnisotropicSimulator", "total_variation_2d", "autotune", "configs",
    "CFRPParameterInverterRicker", "SpatiallyVaryingParameterInverter", "TraceInversion",
    "add_noise_rms", "add_noise_snr", "data", "fem", "FusedTVAdam",
]
