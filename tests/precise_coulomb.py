"""Extended-precision Ewald potential of a unit charge at grid nodes (test infrastructure): the
real-space sum kC sum_j q_j erfc(alpha r_j)/r_j over minimum images within the cutoff in 40-digit
mpmath, and the reciprocal sum kC (4 pi/V) sum_k w(k) Re(conj S(k) exp(i k.r)) over the reference
k-set in 64-bit-mantissa long double.  It pins the Coulomb grid at nodes where the potential is a
cancellation of its two Ewald parts, where any FP64 evaluation carries ~1e-10 K of rounding."""
from __future__ import annotations

import mpmath
import numpy as np

import oracle_lib as ol


class PreciseCoulomb:
    def __init__(self, fw: ol.Framework, alpha: float, kmax: int, rc: float):
        self.fw, self.alpha, self.rc = fw, alpha, rc
        st = ol.EwaldState(fw, alpha, kmax)
        self.kv, self.w, _ = st.kvectors()
        L = np.longdouble
        pos = np.stack([fw.x, fw.y, fw.z], 1).astype(L)
        kr = self.kv.astype(L) @ pos.T
        q = fw.q.astype(L)
        self.sre = (np.cos(kr) * q).sum(1)
        self.sim = (np.sin(kr) * q).sum(1)
        self.pref = L(ol.KCOULOMB) * L(4) * L(np.pi) / L(abs(np.linalg.det(fw.m)))

    def at(self, r):
        """(real, reciprocal) at Cartesian position r, as mpmath numbers."""
        mpmath.mp.dps = 40
        fw = self.fw
        xyz = np.stack([fw.x, fw.y, fw.z], 1)
        d = (xyz - r) @ np.linalg.inv(fw.m)
        d -= np.round(d)
        dist = np.sqrt(((d @ fw.m) ** 2).sum(1))
        sel = dist <= self.rc
        a = mpmath.mpf(self.alpha)
        real = mpmath.mpf(ol.KCOULOMB) * mpmath.fsum(
            mpmath.mpf(float(q)) * mpmath.erfc(a * mpmath.mpf(float(x))) / mpmath.mpf(float(x))
            for q, x in zip(fw.q[sel], dist[sel]))
        L = np.longdouble
        krn = self.kv.astype(L) @ np.asarray(r, dtype=L)
        rec = self.pref * (self.w.astype(L) * (self.sre * np.cos(krn) + self.sim * np.sin(krn))).sum()
        return real, mpmath.mpf(str(rec))

    def node_position(self, grid_m, counts, node):
        nx, ny, nz = counts
        f = np.array([node % nx / nx, (node // nx) % ny / ny, node // (nx * ny) / nz])
        return f @ np.asarray(grid_m)
