"""Deterministic synthetic vector sets for the five benchmark workloads.

Every generator is a pure function of (n, dim, seed) using numpy's PCG64
``default_rng``.  Base rows and queries come from one stream: the mixture
parameters are drawn first, then the base rows, then the queries, so a
smaller ``n`` does not change the mixture.  Rows are produced in chunks so a
100M-row set never holds more than one chunk of float64 temporaries.

Workloads (vector counts, dims, metrics and mixture shapes):
  * gaussian mixture, rounded float32    64 centres U[0,128], sigma 12, clip >= 0, round
  * gaussian mixture, quantised uint8    same mixture, clipped to [0,255] and rounded
  * von Mises-Fisher-like unit vectors   centres on the sphere, normalised noisy copies
  * gaussian mixture, float32            unrounded (traversal copies in fp16 / int8)
"""
from __future__ import annotations

import numpy as np

_CHUNK = 1 << 16


def _mixture_chunks(rng, centers, n, sigma, out_dtype, lo=None, hi=None, rounded=False, rows=1 << 22):
    """The rows of _mixture_rows as consecutive arrays of `rows` rows (a
    multiple of the generation chunk, so the draws are the same)."""
    assert rows % _CHUNK == 0
    k, dim = centers.shape
    for c0 in range(0, n, rows):
        c1 = min(n, c0 + rows)
        out = np.empty((c1 - c0, dim), dtype=out_dtype)
        for s in range(c0, c1, _CHUNK):
            e = min(c1, s + _CHUNK)
            assign = rng.integers(0, k, size=e - s)
            x = centers[assign] + rng.normal(0.0, sigma, size=(e - s, dim))
            if lo is not None or hi is not None:
                x = np.clip(x, lo, hi)
            if rounded:
                x = np.rint(x)
            out[s - c0:e - c0] = x.astype(out_dtype)
        yield out


def _mixture_rows(rng, centers, n, sigma, out_dtype, lo=None, hi=None, rounded=False):
    k, dim = centers.shape
    out = np.empty((n, dim), dtype=out_dtype)
    at = 0
    for c in _mixture_chunks(rng, centers, n, sigma, out_dtype, lo, hi, rounded):
        out[at:at + c.shape[0]] = c
        at += c.shape[0]
    return out


def _mixture_kind(kind):
    if kind == "rounded_f32":
        return dict(out_dtype=np.float32, lo=0.0, hi=None, rounded=True)
    if kind == "u8":
        return dict(out_dtype=np.uint8, lo=0.0, hi=255.0, rounded=True)
    if kind == "f32":
        return dict(out_dtype=np.float32)
    raise ValueError(kind)


class MixtureStream:
    """gaussian_mixture's base rows as a stream of arrays (for sets too large
    to generate before use), then its queries: the same draws in the same
    order, so the rows and queries equal gaussian_mixture's."""

    def __init__(self, n: int, nq: int, dim: int = 128, n_clusters: int = 64, seed: int = 0,
                 sigma: float = 12.0, kind: str = "rounded_f32", rows: int = 1 << 22):
        self.n, self.nq, self.sigma, self.rows = n, nq, sigma, rows
        self.args = _mixture_kind(kind)
        self.rng = np.random.default_rng(seed)
        self.centers = self.rng.uniform(0.0, 128.0, size=(n_clusters, dim))
        self._done = False

    def base_chunks(self):
        for c in _mixture_chunks(self.rng, self.centers, self.n, self.sigma, rows=self.rows, **self.args):
            yield c
        self._done = True

    def queries(self) -> np.ndarray:
        if not self._done:
            raise RuntimeError("draw every base chunk first")
        return _mixture_rows(self.rng, self.centers, self.nq, self.sigma, **self.args)


def gaussian_mixture(n: int, nq: int, dim: int = 128, n_clusters: int = 64, seed: int = 0,
                     sigma: float = 12.0, kind: str = "rounded_f32"):
    """Base (n x dim) and queries (nq x dim) from one Gaussian mixture.

    kind: "rounded_f32" (clip >= 0, round, float32), "u8" (clip to [0,255],
    round, uint8) or "f32" (float32, no clipping or rounding).
    """
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 128.0, size=(n_clusters, dim))
    args = _mixture_kind(kind)
    base = _mixture_rows(rng, centers, n, sigma, **args)
    queries = _mixture_rows(rng, centers, nq, sigma, **args)
    return base, queries


def vmf_mixture(n: int, nq: int, dim: int = 96, n_clusters: int = 256, seed: int = 2,
                spread: float = 0.35):
    """Unit-norm float32 rows around random directions (a vMF-like mixture)."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(size=(n_clusters, dim))
    centers /= np.linalg.norm(centers, axis=1, keepdims=True)

    def rows(m):
        out = np.empty((m, dim), dtype=np.float32)
        for s in range(0, m, _CHUNK):
            e = min(m, s + _CHUNK)
            assign = rng.integers(0, n_clusters, size=e - s)
            x = centers[assign] + rng.normal(0.0, spread / np.sqrt(dim), size=(e - s, dim))
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            out[s:e] = x.astype(np.float32)
        return out

    return rows(n), rows(nq)


def workload(name: str, n: int | None = None, nq: int | None = None):
    """(base, queries, params) for a named workload; n / nq override sizes."""
    if name == "gm100k_f32":
        b, q = gaussian_mixture(n or 100_000, nq or 1_000, 128, 64, 0, kind="rounded_f32")
        return b, q, dict(metric="l2", M=16, efc=200)
    if name == "gm1m_u8":
        b, q = gaussian_mixture(n or 1_000_000, nq or 10_000, 128, 64, 1, kind="u8")
        return b, q, dict(metric="l2", M=32, efc=256)
    if name == "vmf1m_f32":
        b, q = vmf_mixture(n or 1_000_000, nq or 50_000, 96, 256, 2)
        return b, q, dict(metric="ip", M=24, efc=200)
    if name == "gm10m_f32":
        b, q = gaussian_mixture(n or 10_000_000, nq or 100_000, 128, 64, 3, kind="f32")
        return b, q, dict(metric="l2", M=32, efc=200)
    if name == "gm100m_u8":
        b, q = gaussian_mixture(n or 100_000_000, nq or 100_000, 128, 1024, 4, kind="u8")
        return b, q, dict(metric="l2", M=32, efc=256)
    raise ValueError(name)


def workload_stream(name: str, n: int | None = None, nq: int | None = None, rows: int = 1 << 22):
    """(MixtureStream, params) for the Gaussian-mixture workloads: the rows of
    workload(name, n, nq) in chunks, then its queries."""
    spec = {"gm100k_f32": (100_000, 1_000, 64, 0, "rounded_f32", dict(metric="l2", M=16, efc=200)),
            "gm1m_u8": (1_000_000, 10_000, 64, 1, "u8", dict(metric="l2", M=32, efc=256)),
            "gm10m_f32": (10_000_000, 100_000, 64, 3, "f32", dict(metric="l2", M=32, efc=200)),
            "gm100m_u8": (100_000_000, 100_000, 1024, 4, "u8", dict(metric="l2", M=32, efc=256))}
    if name not in spec:
        raise ValueError(f"{name} has no streamed form")
    n0, nq0, k, seed, kind, params = spec[name]
    return MixtureStream(n or n0, nq or nq0, 128, k, seed, kind=kind, rows=rows), params
