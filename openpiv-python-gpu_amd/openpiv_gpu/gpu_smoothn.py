"""smoothn settings of the PIV hot path.

The unweighted, automatic-s DCT-PLS smoother the reference runs between passes
(gpu_smoothn.py:123-372) is a native stage of the pass loop here; this module
keeps the settings constants of the reference module.
"""

SMOOTHING_WEIGHTS = None
INIT_GUESS = None
SPACING = None
SMOOTHING_PAR = None
SMOOTHING_ORDER = 2
MAX_ROBUST_STEPS = 3
MAX_SMOOTHING_ITERS = 100
TOL = 1e-3
ROBUST = False
WEIGHTING_METHOD = "bisquare"
N_P0 = 10
COARSE_COEFFICIENTS = 10
ERR_P = 0.1
