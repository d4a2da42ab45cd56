"""MI355X-native Landau-de Gennes Q-tensor relaxation / quench engine.

The compute path is libqsr_amd.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/qsr_amd.h); :mod:`lcdrsimulator_amd.qsr` is its ctypes
binding and ``lib/QSR_hip`` the ``--config`` command-line driver.
"""
from .qsr import (PROTOCOL_CONSTANT, PROTOCOL_RAMP, PROTOCOL_STEP, SURR_CHECKED, SURR_FORCE, SURR_PROPOSE,
                  SURR_INPUT_FIELD, SURR_INPUT_INCREMENT, SURR_INPUT_GAIN, SURR_INPUT_STEP, SURR_TILES_ALL,
                  SURR_TILES_DROPLET,
                  Params,  # noqa: F401
                  QSRError, Solver, S_eq, Surrogate, device_count, dt_estimate, init_field, load, load_qtensor, save_qtensor)
