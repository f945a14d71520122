"""Calibration probe for HBM byte counters: copies a known number of bytes on the GPU.

Run under `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE): the copy kernel reads and
writes exactly NBYTES each, so the counter's reading over NBYTES is the scale
factor of that counter on this device.
"""
import torch

NBYTES = 512 * 1024 * 1024


def main():
    n = NBYTES // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {NBYTES} bytes x 3", flush=True)


if __name__ == "__main__":
    main()
