"""Plain device-to-device copy of a known size, the calibration run for HBM byte counters.

    rocprofv3 --pmc FETCH_SIZE -- python -m yapt_amd.calib_copy [MiB]

FETCH_SIZE / WRITE_SIZE of this copy against its known size gives the factor that turns the
counters of the path-tracing kernels into bytes.
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copied %d MiB three times" % mib)


if __name__ == "__main__":
    main()
