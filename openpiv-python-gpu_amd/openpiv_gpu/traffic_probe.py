"""Workload for the HBM-traffic counter runs: the 2048x2048 uint16 64->32->16
pass loop on a device-resident pair, plus device-to-device copies of a known
size that calibrate the FETCH_SIZE / WRITE_SIZE counters.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o fetch -- python -m openpiv_gpu.traffic_probe
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -o write -- python -m openpiv_gpu.traffic_probe
"""

import sys

import numpy as np

from . import _native, synthetic
from .gpu_process import PIVGPU

COPY_BYTES = 256 << 20


def main(steps=3, copies=3):
    shape = (2048, 2048)
    gen = synthetic.DevicePairGenerator(shape, synthetic.MODEL_12BIT)
    fa, fb = gen.render(synthetic.sine_displacement(4.0, 3.0, 256.0), seed=2)
    piv = PIVGPU(shape, 16, search_size_iters=(1, 1, 1), overlap_ratio=0.5, median_tol=2, dtype_f="float32",
                 use_graph=False)
    plan = piv._get_plan(_native.PIV_U16)
    m, n = piv.piv_fields[-1].field_shape
    u = _native.DeviceBuffer((m, n), np.float32)
    v = _native.DeviceBuffer((m, n), np.float32)
    for _ in range(steps):
        plan.run(fa.value, fb.value, True, u.value, v.value, None, None, True)
    src = _native.DeviceBuffer((COPY_BYTES,), np.uint8)
    dst = _native.DeviceBuffer((COPY_BYTES,), np.uint8)
    lib = _native.load_library()
    for _ in range(copies):
        _native.check(lib.piv_memcpy(dst.ptr, src.ptr, COPY_BYTES, 2))
    _native.synchronize()
    print(f"probe done: {steps} pass loops, {copies} copies of {COPY_BYTES} bytes", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
