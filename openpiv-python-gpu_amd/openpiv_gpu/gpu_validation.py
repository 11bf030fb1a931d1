"""Validation and replacement settings of the PIV hot path.

The reference exposes ``ValidationGPU`` / ``ReplacementGPU`` objects
(gpu_validation.py:25-415) that ``PIVGPU`` builds every pass.  Here the median
test and the spring replacement run inside the native pass loop; this module
keeps the settings constants the drop-in API imports.
"""

# Default validation settings (gpu_validation.py:10-16).
VALIDATION_SIZE = 1
S2N_TOL = None
MEDIAN_TOL = 2
MAD_TOL = None
MEAN_TOL = None
RMS_TOL = None

# Default replacement settings (gpu_validation.py:18-23).
REPLACING_METHOD = "spring"
REPLACING_SIZE = 1
ALLOWED_REPLACING_METHODS = {"spring", "median", "mean"}
