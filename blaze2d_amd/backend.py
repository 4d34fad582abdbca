"""SpectralBackend and Theta operator on device buffers (C ABI).

Mirrors ``trait SpectralBackend`` (crates/core/src/backend.rs:73-165), the
``CpuBackend`` it replaces (crates/backend-cpu/src/lib.rs:28-452) and the
``LinearOperator`` / ``OperatorPreconditioner`` traits
(crates/core/src/operators.rs:91-165, crates/core/src/preconditioners.rs:102-113).
Sizes are checked on the host before any call; errors raise RuntimeError.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib as L


class HipBackend:
    """One device context (a HIP stream)."""

    def __init__(self, device: int = 0):
        self.device = device
        self._ctx = ctypes.c_void_p()
        L.check(L.lib().bh_context_create(device, ctypes.byref(self._ctx)), "bh_context_create")

    @property
    def handle(self):
        return self._ctx

    def close(self):
        if self._ctx:
            L.lib().bh_context_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- SpectralBackend -----------------------------------------------------
    def alloc_field(self, nx: int, ny: int, count: int = 1, precision: int = 0) -> "Field":
        return Field(self, nx, ny, count, precision)

    def forward_fft_2d(self, buf: "Field", first: int = 0, count: Optional[int] = None):
        count = buf.count - first if count is None else count
        L.check(L.lib().bh_fft2d(self._ctx, buf.handle, first, count, -1), "forward_fft_2d")

    def inverse_fft_2d(self, buf: "Field", first: int = 0, count: Optional[int] = None):
        count = buf.count - first if count is None else count
        L.check(L.lib().bh_fft2d(self._ctx, buf.handle, first, count, 1), "inverse_fft_2d")

    def scale(self, alpha: complex, buf: "Field", index: int = 0):
        L.check(L.lib().bh_scale(self._ctx, buf.handle, index, alpha.real, alpha.imag), "scale")

    def axpy(self, alpha: complex, x: "Field", y: "Field", ix: int = 0, iy: int = 0):
        L.check(L.lib().bh_axpy(self._ctx, alpha.real, alpha.imag, x.handle, ix, y.handle, iy), "axpy")

    def dot(self, x: "Field", y: "Field", ix: int = 0, iy: int = 0) -> complex:
        out = np.zeros(2)
        L.check(L.lib().bh_dot(self._ctx, x.handle, ix, y.handle, iy, out.ctypes.data_as(L.c_dp)), "dot")
        return complex(out[0], out[1])

    def gram_matrix(self, x: "Field", y: "Field", xi0: int = 0, p: Optional[int] = None, yi0: int = 0,
                    q: Optional[int] = None) -> np.ndarray:
        p = x.count - xi0 if p is None else p
        q = y.count - yi0 if q is None else q
        out = np.zeros((p, q), dtype=np.complex128)
        L.check(L.lib().bh_gram(self._ctx, x.handle, xi0, p, y.handle, yi0, q,
                                out.view(np.float64).ctypes.data_as(L.c_dp)), "gram_matrix")
        return out

    def hermitian_eigh(self, mats: np.ndarray):
        """Batched dense Hermitian eigendecomposition (ascending), FP64 on the device."""
        a = np.asarray(mats, dtype=np.complex128)
        if a.ndim == 2:
            a = a[None]
        cnt, n, _ = a.shape
        flat = np.ascontiguousarray(np.transpose(a, (0, 2, 1)).reshape(cnt, -1))  # col-major per matrix
        w = np.zeros((cnt, n))
        v = np.zeros((cnt, n * n), dtype=np.complex128)
        L.check(L.lib().bh_hermitian_eigh(self._ctx, n, cnt, flat.view(np.float64).ctypes.data_as(L.c_dp),
                                          w.ctypes.data_as(L.c_dp), v.view(np.float64).ctypes.data_as(L.c_dp)),
                "hermitian_eigh")
        return w, np.transpose(v.reshape(cnt, n, n), (0, 2, 1))

    def rayleigh_ritz(self, g1: np.ndarray, g2: np.ndarray, nsplit: int):
        """Batched Rayleigh-Ritz from the Grams G1 = Z^H B Z, G2 = Z^H A Z ([count,] n, n).

        Returns (lam[count, n], coef[count, n, n], info[count, 4]) as
        bh_rayleigh_ritz: ascending Ritz values (rank of them), coefficients of
        the new X and W Ritz vectors over Z's columns, and {rank, SVQB path
        taken, cnt_x, cnt_w}."""
        a1 = np.asarray(g1, dtype=np.complex128)
        a2 = np.asarray(g2, dtype=np.complex128)
        if a1.ndim == 2:
            a1, a2 = a1[None], a2[None]
        cnt, n, _ = a1.shape
        f1 = np.ascontiguousarray(np.transpose(a1, (0, 2, 1)).reshape(cnt, -1))
        f2 = np.ascontiguousarray(np.transpose(a2, (0, 2, 1)).reshape(cnt, -1))
        lam = np.zeros((cnt, n))
        coef = np.zeros((cnt, n * n), dtype=np.complex128)
        info = np.zeros((cnt, 4), dtype=np.int32)
        L.check(L.lib().bh_rayleigh_ritz(self._ctx, n, cnt, nsplit, f1.view(np.float64).ctypes.data_as(L.c_dp),
                                         f2.view(np.float64).ctypes.data_as(L.c_dp), lam.ctypes.data_as(L.c_dp),
                                         coef.view(np.float64).ctypes.data_as(L.c_dp),
                                         info.ctypes.data_as(L.c_ip)), "rayleigh_ritz")
        return lam, np.transpose(coef.reshape(cnt, n, n), (0, 2, 1)), info

    def linear_combinations(self, q: "Field", coeffs: np.ndarray, out: "Field", qi0: int = 0, oi0: int = 0):
        c = np.asfortranarray(coeffs, dtype=np.complex128)
        r, m = c.shape
        flat = np.ascontiguousarray(c.reshape(-1, order="F"))
        L.check(L.lib().bh_lincomb(self._ctx, q.handle, qi0, r, flat.view(np.float64).ctypes.data_as(L.c_dp), m,
                                   out.handle, oi0), "linear_combinations")


class Field:
    """`count` device fields of nx*ny complex values (Field2D, field.rs:101-162)."""

    def __init__(self, backend: HipBackend, nx: int, ny: int, count: int = 1, precision: int = 0):
        self.backend = backend
        self.nx, self.ny, self.count, self.precision = nx, ny, count, precision
        self._h = ctypes.c_void_p()
        L.check(L.lib().bh_buffer_alloc(backend.handle, nx, ny, precision, count, ctypes.byref(self._h)),
                "bh_buffer_alloc")

    @property
    def handle(self):
        return self._h

    def upload(self, data: np.ndarray, first: int = 0):
        a = np.ascontiguousarray(data, dtype=np.complex128).reshape(-1, self.nx * self.ny)
        if first + a.shape[0] > self.count:
            raise ValueError("upload beyond the buffer")
        L.check(L.lib().bh_buffer_upload(self._h, first, a.shape[0], a.view(np.float64).ctypes.data_as(L.c_dp)),
                "upload")

    def download(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = self.count - first if count is None else count
        out = np.zeros((count, self.nx * self.ny), dtype=np.complex128)
        L.check(L.lib().bh_buffer_download(self._h, first, count, out.view(np.float64).ctypes.data_as(L.c_dp)),
                "download")
        return out

    def __del__(self):
        try:
            if self._h:
                L.lib().bh_buffer_free(self._h)
        except Exception:
            pass


class ThetaOperator:
    """Theta (TE/TM), mass and preconditioner at one k-point (maxwell.rs:126-379)."""

    def __init__(self, backend: HipBackend, nx: int, ny: int, pol: str, b1, b2, bloch, eps: np.ndarray,
                 tensor: Optional[np.ndarray] = None, precond: int = -1, precision: int = 0):
        self.backend = backend
        self.nx, self.ny = nx, ny
        self._eps = np.ascontiguousarray(eps, dtype=np.float64)
        self._ten = None if tensor is None else np.ascontiguousarray(tensor, dtype=np.float64)
        if self._eps.size != nx * ny or (self._ten is not None and self._ten.size != 4 * nx * ny):
            raise ValueError("table sizes do not match the grid")
        d = L.ThetaDesc()
        d.nx, d.ny = nx, ny
        d.pol = 0 if pol.upper() == "TM" else 1
        for i in range(2):
            d.b1[i], d.b2[i], d.bloch[i] = b1[i], b2[i], bloch[i]
        d.eps = self._eps.ctypes.data_as(L.c_dp)
        d.tensor = self._ten.ctypes.data_as(L.c_dp) if self._ten is not None else None
        d.precond = precond
        d.precision = precision
        self._h = ctypes.c_void_p()
        L.check(L.lib().bh_theta_create(backend.handle, ctypes.byref(d), ctypes.byref(self._h)), "bh_theta_create")

    def apply(self, inp: Field, out: Field, ii0: int = 0, oi0: int = 0, count: Optional[int] = None):
        count = inp.count - ii0 if count is None else count
        L.check(L.lib().bh_theta_apply(self._h, inp.handle, ii0, out.handle, oi0, count), "theta_apply")

    def apply_mass(self, inp: Field, out: Field, ii0: int = 0, oi0: int = 0, count: Optional[int] = None):
        count = inp.count - ii0 if count is None else count
        L.check(L.lib().bh_mass_apply(self._h, inp.handle, ii0, out.handle, oi0, count), "mass_apply")

    def precondition(self, buf: Field, first: int = 0, count: Optional[int] = None):
        count = buf.count - first if count is None else count
        L.check(L.lib().bh_precond_apply(self._h, buf.handle, first, count), "precond_apply")

    def __del__(self):
        try:
            if self._h:
                L.lib().bh_theta_destroy(self._h)
        except Exception:
            pass
