"""Device-memory plumbing for the GPU tests (torch allocates, the C ABI computes)."""
import ctypes as C

import numpy as np
import torch

import hedgehog_lpbf_amd as h


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def stream():
    return torch.cuda.current_stream().cuda_stream


def sync_check():
    torch.cuda.synchronize()
    h.check()


def kconst(k):
    """Copy an oracle StepConst into the package's identical ctypes record."""
    return h.StepConst.from_buffer_copy(bytes(k))


def plan_dev(plan: np.ndarray):
    return dev(np.ascontiguousarray(plan, dtype=np.float64))


def stats_np(buf: torch.Tensor, n: int):
    raw = bytes(buf.cpu().numpy().tobytes())
    arr = (h.StepStats * n).from_buffer_copy(raw[: n * C.sizeof(h.StepStats)])
    return list(arr)


def partials_np(buf: torch.Tensor, n: int):
    raw = bytes(buf.cpu().numpy().tobytes())
    return list((h.Partial * n).from_buffer_copy(raw[: n * C.sizeof(h.Partial)]))


def dev_bytes(nbytes: int):
    return torch.empty(max(nbytes, 8) // 8 + 1, dtype=torch.float64, device="cuda")
