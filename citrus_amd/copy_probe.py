"""Plain device copy of known size, the calibration for HBM byte counters.

``python -m citrus_amd.copy_probe [MiB]`` copies a float32 buffer of that size (default 1024 MiB) three times on
cuda:0. Under ``rocprofv3 --pmc FETCH_SIZE`` (then ``WRITE_SIZE``) each copy dispatch moves exactly the buffer size in
each direction, so the counters' per-dispatch readings give the factors that turn them into bytes for other kernels.
"""

from __future__ import annotations

import sys


def main() -> None:
    import torch

    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda:0", dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy_probe: 3 copies of {mib} MiB ({n * 4} bytes each way)", flush=True)


if __name__ == "__main__":
    main()
