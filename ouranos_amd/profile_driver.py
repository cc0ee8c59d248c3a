"""Minimal driver for hardware-counter runs of the sweep.

One operator application of a named workload on device buffers, followed by
a 1 GiB device copy through the library's plain copy kernel (the reference
point that scales the FETCH_SIZE / WRITE_SIZE traffic counters).

    python -m ouranos_amd.profile_driver [workload] [fp32|fp64] [levels|persistent|tiles]

Without a mode the plan's default runs (tiles for 2D Q1-Q2 plans that have
them, level launches otherwise).
"""
import ctypes
import sys

import ouranos_amd as ou
from ouranos_amd import problems as P


def main(argv):
    name = argv[1] if len(argv) > 1 else "headline_3d"
    prec = ou.FP64 if len(argv) > 2 and argv[2] == "fp64" else ou.FP32
    spec = getattr(P, name)()
    op = spec.operator(prec)
    if len(argv) > 3:
        op.sweep_mode = {"levels": ou.SWEEP_LEVELS, "persistent": ou.SWEEP_PERSISTENT,
                         "tiles": ou.SWEEP_TILES}[argv[3]]
    x = P.seeded_vector(spec.size, 1)
    d_x = ou.DeviceBuffer(x.nbytes)
    d_y = ou.DeviceBuffer(x.nbytes)
    d_x.upload(x)
    op.time_apply(d_x, d_y, 0, 1)
    n = 1 << 30
    a, b = ou.DeviceBuffer(n), ou.DeviceBuffer(n)
    lib = ou.lib()
    lib.ou_memcpy_dtod.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.ou_memcpy_dtod(b.ptr, a.ptr, n)
    lib.ou_device_synchronize()
    print("profile driver done", flush=True)


if __name__ == "__main__":
    main(sys.argv)
