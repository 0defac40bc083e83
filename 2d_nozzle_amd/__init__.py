"""MI355X-native DG residual and SSP-RK3 engine for the 2D nozzle Euler solver.

The package directory name starts with a digit, so import it with
``importlib.import_module("2d_nozzle_amd")`` (or load it under an alias).
"""

__version__ = "0.1.0"
