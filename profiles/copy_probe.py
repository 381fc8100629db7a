"""Calibration workload for the rocprofv3 HBM-traffic counters: a plain
device-to-device copy of a known size (torch on the GPU), so FETCH_SIZE /
WRITE_SIZE per dispatch can be scaled to bytes.  Usage:
  rocprofv3 --pmc FETCH_SIZE -- python3 profiles/copy_probe.py
"""
import torch

n = 256 * 1024 * 1024 // 4          # 256 MiB of float32
a = torch.ones(n, dtype=torch.float32, device="cuda")
b = torch.empty_like(a)
for _ in range(5):
    b.copy_(a)
torch.cuda.synchronize()
print("copied %d bytes x5" % (4 * n))
