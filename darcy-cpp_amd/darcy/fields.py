"""Random permeability fields (the inputs of every benchmark configuration).

Same sampling procedure as the reference's ``python/darcy/fields.py:31-83``:
a mean-zero Gaussian random field with spectral density
``tau^(alpha-1) (pi^2 (k1^2 + k2^2) + tau^2)^(-alpha/2)`` (covariance
``(-Laplacian + tau^2 I)^(-alpha)``), drawn by scaling i.i.d. normals from
``numpy.random.default_rng(seed).standard_normal((n, n))`` and taking the
orthonormal 2-D inverse DCT (DCT-III along both axes), then thresholded at
zero into a two-phase medium.  The steps have to stay identical to the
reference's, draw for draw, or the media -- and therefore the published
iteration counts -- would change.

The inverse DCT uses ``scipy.fft.idctn(norm="ortho")`` when SciPy is present
(as the reference does); otherwise an FFT-based DCT-III written here, which
agrees with SciPy to round-off (checked in the tests).
"""

from __future__ import annotations

from typing import Union

import numpy as np

__all__ = ["gaussian_random_field", "two_phase_permeability", "dct3_ortho"]

SeedLike = Union[int, np.random.Generator, None]


def _generator(seed: SeedLike) -> np.random.Generator:
    if isinstance(seed, np.random.Generator):
        return seed
    return np.random.default_rng(seed)


def _dct3_axis(x: np.ndarray, axis: int) -> np.ndarray:
    """Orthonormal DCT-III along one axis via a length-2N complex FFT.

    y[m] = x[0]/sqrt(N) + sqrt(2/N) * sum_{k>=1} x[k] cos(pi k (2m+1) / (2N))
         = Re( sum_k c_k exp(i pi k / (2N)) exp(2 pi i k m / (2N)) ).
    """
    x = np.moveaxis(x, axis, -1)
    n = x.shape[-1]
    k = np.arange(n)
    w = np.full(n, np.sqrt(2.0 / n))
    w[0] = np.sqrt(1.0 / n)
    c = x * (w * np.exp(1j * np.pi * k / (2.0 * n)))
    padded = np.zeros(x.shape[:-1] + (2 * n,), dtype=np.complex128)
    padded[..., :n] = c
    y = np.fft.ifft(padded, axis=-1)[..., :n].real * (2.0 * n)
    return np.moveaxis(y, -1, axis)


def dct3_ortho(x: np.ndarray) -> np.ndarray:
    """2-D orthonormal inverse DCT (scipy.fft.idctn(x, norm='ortho'))."""
    return _dct3_axis(_dct3_axis(np.asarray(x, dtype=np.float64), 0), 1)


def _idctn_ortho(x: np.ndarray) -> np.ndarray:
    try:
        from scipy.fft import idctn
    except ImportError:  # pragma: no cover - SciPy is in the image
        return dct3_ortho(x)
    return np.asarray(idctn(x, norm="ortho"), dtype=np.float64)


def gaussian_random_field(
    n: int,
    alpha: float = 2.0,
    tau: float = 3.0,
    seed: SeedLike = None,
    rng: np.random.Generator | None = None,
) -> np.ndarray:
    """Sample a mean-zero Gaussian random field on an ``n x n`` grid.

    Follows python/darcy/fields.py:31-69 of the reference: ``rng`` is the
    deprecated alias for ``seed``.
    """
    if n < 1:
        raise ValueError(f"n must be positive, got {n}")
    gen = _generator(rng if rng is not None else seed)
    wave = np.arange(n)
    k1, k2 = np.meshgrid(wave, wave, indexing="ij")
    spectrum = (np.pi**2) * (k1**2 + k2**2) + tau**2
    amplitude = (tau ** (alpha - 1.0)) * spectrum ** (-alpha / 2.0)
    amplitude[0, 0] = 0.0  # the mean mode is dropped: a centred field
    xi = gen.standard_normal((n, n))
    return _idctn_ortho(amplitude * xi)


def two_phase_permeability(grf: np.ndarray, hi: float = 12.0, lo: float = 3.0) -> np.ndarray:
    """``hi`` where the field is >= 0, ``lo`` elsewhere (fields.py:72-83)."""
    if hi <= 0 or lo <= 0:
        raise ValueError("permeability phases must be strictly positive")
    return np.where(np.asarray(grf) >= 0.0, hi, lo).astype(np.float64)
