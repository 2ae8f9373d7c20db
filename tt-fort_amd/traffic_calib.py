"""Known-size device copies for calibrating the HBM traffic counters.

rocprofv3's FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are read for a plain
device-to-device copy of a known size; their ratio to the true byte count is
the scale factor applied to the same counters of the measured kernel.

    rocprofv3 --pmc FETCH_SIZE -- python3 tt-fort_amd/traffic_calib.py
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import ttfort  # noqa: E402


def main(mib=256, copies=3):
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    nbytes = mib << 20
    a = ttfort.DeviceArray((nbytes // 8,))
    b = ttfort.DeviceArray((nbytes // 8,))
    ttfort.check(lib.ttf_memset(a.ptr, 1, nbytes), lib)
    for _ in range(copies):
        ttfort.check(lib.ttf_memcpy_d2d(b.ptr, a.ptr, nbytes), lib)
    ttfort.sync(lib)
    print(json.dumps({"copy_bytes": nbytes, "copies": copies}))


if __name__ == "__main__":
    main()
