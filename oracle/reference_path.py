"""The reference CPU hot path driven through the reference's own C++ interpolator.

Test infrastructure and CPU baseline only.  ``ReferenceGridPath`` restates
the Python around ``ginterp`` call for call -- HighPerfInterpolator.interpolate
(interpolate.py:475-543: ``np.empty_like`` output, per-component buffers),
Interp2D/Interp3D.__call__ (grid_interp.py:69-85,155-172: physical -> index
conversion per component call) and RungeKutta4Integrator.integrate
(integrate.py:291-342: k1..k4, tmp, ``np.copyto``) -- and calls the compiled
reference ``GridInterp.cpp`` (``oracle/_ref``) for the interpolation itself.
The only substitution is NumPy for the one Numba line of the RK4 update
(integrate.py:81-109; Numba is not installed here).
"""

from __future__ import annotations

import numpy as np

from . import native


class ReferenceGridPath:
    def __init__(self, v_components, mins, deltas):
        ref = native.reference_ginterp()
        if ref is None:
            raise RuntimeError("oracle/_ref is not built")
        self.ref = ref
        self.v = [np.ascontiguousarray(c, dtype=np.float64) for c in v_components]
        self.dim = len(self.v)
        self.mins = list(mins)
        self.deltas = list(deltas)
        self._buf = None
        self._ub = None

    def _component(self, c, t, out):
        # grid_interp.py:69-85 / 155-172
        arr = np.atleast_2d(t)
        n = arr.shape[0]
        if self._buf is None or self._buf.shape[0] != n:
            self._buf = np.empty_like(arr)
        points = self._buf
        for a in range(self.dim):
            points[:, a] = (arr[:, a] - self.mins[a]) / self.deltas[a]
        if self.dim == 2:
            self.ref.interp2d_vec_inplace(self.v[c], points, out)
        else:
            self.ref.interp3d_vec_inplace(self.v[c], points, out)
        return out

    def interpolate(self, new_points):
        # interpolate.py:475-543
        n = new_points.shape[0]
        out = np.empty_like(new_points)
        if self._ub is None or self._ub[0].shape[0] != n:
            self._ub = [np.empty(n) for _ in range(self.dim)]
        for c in range(self.dim):
            self._component(c, new_points, self._ub[c])
            out[:, c] = self._ub[c]
        return out

    def rk4_steps(self, positions, h, n_steps):
        # integrate.py:291-342
        k1 = np.empty_like(positions)
        k2 = np.empty_like(positions)
        k3 = np.empty_like(positions)
        k4 = np.empty_like(positions)
        tmp = np.empty_like(positions)
        for _ in range(n_steps):
            np.copyto(k1, self.interpolate(positions))
            np.copyto(tmp, positions)
            tmp += 0.5 * h * k1
            np.copyto(k2, self.interpolate(tmp))
            np.copyto(tmp, positions)
            tmp += 0.5 * h * k2
            np.copyto(k3, self.interpolate(tmp))
            np.copyto(tmp, positions)
            tmp += h * k3
            np.copyto(k4, self.interpolate(tmp))
            positions += (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        return positions
