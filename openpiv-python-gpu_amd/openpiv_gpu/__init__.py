"""MI355X-native drop-in for the openpiv-python-gpu PIV hot path.

``from openpiv_gpu.gpu_process import piv_gpu`` works as with the reference;
the pass loop runs in HIP kernels for gfx950 (see gpu_process.py).
"""

import numpy as np

# Default data types (same names as the reference package).
DTYPE_i = np.int64
DTYPE_f = np.float64
DTYPE_c = np.complex128

__all__ = ["DTYPE_i", "DTYPE_f", "DTYPE_c"]
