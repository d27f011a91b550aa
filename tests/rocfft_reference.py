"""Child program of test_gpu_selfgravity: the SG convolution evaluated
with torch.fft on the GPU (rocFFT underneath), in its own process so that
torch's HIP runtime is the only one initialised in it.

usage: python rocfft_reference.py in.npz out.npz
"""
import sys

import numpy as np
import torch


def main(src, dst):
    z = np.load(src)
    dev = torch.device("cuda:0")
    f = lambda a: torch.fft.fft2(torch.tensor(a, dtype=torch.complex128, device=dev))  # noqa: E731
    n = z["Sr"].size
    nr = z["Sr"].shape[0] // 2
    Gr = (torch.fft.ifft2(-f(z["Kr"]) * f(z["Sr"])) * n).real.cpu().numpy()[:nr]
    Gt = (torch.fft.ifft2(-f(z["Kt"]) * f(z["St"])) * n).real.cpu().numpy()[:nr]
    np.savez(dst, Gr=Gr, Gt=Gt)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
