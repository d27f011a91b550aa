"""Plain device-to-device copy of a known size (3 x 1 GiB), the reference
point for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters into HBM
bytes:  rocprofv3 --pmc FETCH_SIZE -- python3 -m cudaprotoplanetarydisks_amd.copyprobe
(run it in a process of its own: it initialises torch's HIP runtime)."""
import torch


def main():
    n = 1 << 27  # 2^27 doubles = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied 3 x {n * 8} bytes")


if __name__ == "__main__":
    main()
