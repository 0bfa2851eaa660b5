"""Helpers for the GPU tests: host <-> device copies through torch and a
device-resident CSR matrix."""
import sys

import numpy as np


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev(a, dtype=None):
    import torch

    a = np.ascontiguousarray(a)
    if dtype is not None:
        a = a.astype(dtype)
    return torch.from_numpy(a).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def empty(n, dtype):
    import torch

    return torch.zeros(max(int(n), 1), dtype=dtype, device="cuda")


class DevCSR:
    def __init__(self, A, f32=False):
        import torch

        self.m, self.n, self.nnz = A.m, A.n, A.nnz
        self.row_ptr = dev(A.row_ptr, np.int32)
        self.col_ind = dev(A.col_ind if A.nnz else np.zeros(1, np.int32), np.int32)
        v = A.val if A.nnz else np.zeros(1)
        self.val = dev(v, np.float32 if f32 else np.float64)
        self.dtype = torch.float32 if f32 else torch.float64


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))
