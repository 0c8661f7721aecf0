"""Range noise on the device: the counterparts of Rmagine's NoiseCuda family
(src/rmagine_cuda/include/rmagine/noise/NoiseCuda.hpp, GaussianNoiseCuda.hpp,
RelGaussianNoiseCuda.hpp, UniformDustNoiseCuda.hpp; kernels in
src/rmagine_cuda/src/noise/*.cu), over rmh_noise_apply.

    noise = GaussianNoiseHip(mean=0.0, stddev=0.01)
    noise.apply(res.ranges)          # in place, float32 CUDA tensor

Options mirror NoiseCuda::Options (seed, max_range); the memory options of the reference
(fixed_memory, estimated_memory_size, never_shrink_memory) have nothing to manage here:
the random numbers are counter-based (Philox4x32-10 of element index, call count and
seed), so no per-element state is kept.  Each apply() uses the next call count, so
successive calls draw new numbers; two objects with the same seed draw the same ones.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from . import types as T


@dataclass
class Options:
    seed: int = 42
    max_range: float = 10000.0


def _stream(stream):
    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(getattr(stream, "cuda_stream", stream))


class NoiseHip:
    kind = -1

    def __init__(self, p0: float, p1: float, p2: float = 0.0, options: Options | None = None):
        self.options = options or Options()
        self._params = T.rmh_noise_params(self.kind, p0, p1, p2, self.options.max_range, self.options.seed & 0xFFFFFFFF)
        self.calls = 0

    def apply(self, ranges: torch.Tensor, stream=None) -> torch.Tensor:
        if not (isinstance(ranges, torch.Tensor) and ranges.dtype == torch.float32 and ranges.is_cuda
                and ranges.is_contiguous()):
            raise ValueError("ranges must be a contiguous float32 CUDA tensor")
        _lib.call("rmh_noise_apply", ctypes.byref(self._params), ctypes.c_void_p(ranges.data_ptr()),
                  ranges.numel(), self.calls & 0xFFFFFFFF, _stream(stream))
        self.calls += 1
        return ranges


class GaussianNoiseHip(NoiseHip):
    """r += N(0,1) * stddev + mean for every r <= max_range (GaussianNoiseCuda.cu:7-27)."""
    kind = T.NOISE_GAUSSIAN

    def __init__(self, mean: float, stddev: float, options: Options | None = None):
        super().__init__(mean, stddev, 0.0, options)


class RelGaussianNoiseHip(NoiseHip):
    """r += N(0,1) * stddev * r^range_exp + mean for every r <= max_range
    (RelGaussianNoiseCuda.cu:7-29)."""
    kind = T.NOISE_REL_GAUSSIAN

    def __init__(self, mean: float, stddev: float, range_exp: float, options: Options | None = None):
        super().__init__(mean, stddev, range_exp, options)


class UniformDustNoiseHip(NoiseHip):
    """Dust along each beam: with probability 1 - (1 - hit_probability)^d (d = min(r,
    max_range)) a particle at u*d is seen and, with probability return_probability^(u d),
    replaces the range (UniformDustNoiseCuda.cu:7-45)."""
    kind = T.NOISE_UNIFORM_DUST

    def __init__(self, hit_probability: float, return_probability: float, options: Options | None = None):
        super().__init__(hit_probability, return_probability, 0.0, options)
