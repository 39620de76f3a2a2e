"""Helpers shared by the test modules."""

import os

import numpy as np
import torch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    path = os.path.join(GOLDEN, name)
    return np.load(path, allow_pickle=False)


def torch_sin32(x):
    """float32 sine from torch's CPU kernel: the arithmetic the reference's
    ``_source_waveform`` uses (fdtd3d.py:333)."""
    return torch.sin(torch.tensor(np.float32(x), dtype=torch.float32)).item()


def rel_l2(a, b):
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
    a = np.asarray(a, dtype=np.complex128 if cplx else np.float64)
    b = np.asarray(b, dtype=np.complex128 if cplx else np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
