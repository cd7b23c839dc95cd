"""Helpers shared by the GPU parity tests."""
import numpy as np
import pytest


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def dev(a, dtype=None):
    torch = torch_cuda()
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)
