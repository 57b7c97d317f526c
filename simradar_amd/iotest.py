"""Bandwidth probe of the engine (RS_io_test): copies n float4 a number of times
with the engine's k_io kernel and prints the read + write rate.  Also the
calibration workload for rocprofv3 FETCH_SIZE / WRITE_SIZE counters: each
dispatch reads and writes exactly 16 n bytes.

    python -m simradar_amd.iotest [n_float4] [reps]
"""
import ctypes
import sys

import simradar_amd as sa


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 8 * 1024 * 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    lib = sa.load()
    e = ctypes.c_void_p()
    if lib.rsd_open(0, ctypes.byref(e)):
        raise RuntimeError("no GPU")
    g = ctypes.c_double()
    rc = lib.rsd_io_test(e, n, reps, ctypes.byref(g))
    if rc:
        raise RuntimeError(lib.rsd_last_error(e).decode())
    print("io: %d float4 (%d bytes read + %d written per dispatch), %d dispatches, %.1f GB/s"
          % (n, 16 * n, 16 * n, reps, g.value))
    lib.rsd_close(e)


if __name__ == "__main__":
    main()
