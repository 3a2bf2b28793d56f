"""Shared helpers for the GPU parity tests (device buffers through torch)."""
import numpy as np


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def empty(n):
    import torch
    return torch.full((int(n),), float("nan"), dtype=torch.float64, device="cuda")


def rel_max(a, b):
    """max |a - b| / max |b| (the per-field metric of the measurement plan)."""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def ulp_diff(a, b):
    """Largest distance in units in the last place between two float64 arrays."""
    ia = np.asarray(a, dtype=np.float64).view(np.int64)
    ib = np.asarray(b, dtype=np.float64).view(np.int64)
    # map to a monotone integer line
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return int(np.abs(ia - ib).max()) if ia.size else 0
