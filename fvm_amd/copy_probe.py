"""Plain device copy of a known size, for calibrating HBM byte counters.

python -m fvm_amd.copy_probe [MiB]: copies a float64 array of the given size (default 512 MiB)
four times with torch's copy kernel; under rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) each copy
dispatch should read (write) exactly that many bytes, which fixes the scale factor applied to
the counters of the solver's kernels.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 512
    n = mib * 1024 * 1024 // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 4", flush=True)


if __name__ == "__main__":
    main()
