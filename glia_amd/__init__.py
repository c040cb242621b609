"""glia_amd: MI355X-native reaction-diffusion forward / adjoint / gradient path.

The compute path is libglia_rd.so (HIP kernels for gfx950 behind the C ABI of include/glia_rd.h).
Python here is a thin binding: rd.py mirrors the reference operator classes, synth.py generates
the synthetic tissue inputs of the benchmark configurations, autograd.py wraps the gradient in a
torch.autograd.Function.
"""
from . import _lib  # noqa: F401

__all__ = ["rd", "synth", "autograd"]
