"""A plain device copy of known size, for calibrating HBM byte counters:
five copies of 512 MiB (torch's elementwise copy kernel), so FETCH_SIZE
and WRITE_SIZE per dispatch should read 524288 KiB each.

  python -m mdbas_amd.copy_probe
"""
import torch


def main():
    n = 512 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 5 x 512 MiB", flush=True)


if __name__ == "__main__":
    main()
