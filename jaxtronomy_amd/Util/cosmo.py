"""Background cosmology for multi-plane lensing distances.

The reference takes an astropy cosmology (default_cosmology, Planck18) through
lenstronomy's Background class (Cosmo/background.py); neither library is
available here.  FlatLambdaCDM restates the distances for a flat universe of
matter and a cosmological constant (no radiation, no massive neutrinos, so
astropy's Planck18 -- which has both -- differs from FlatLambdaCDM(67.66,
0.30966) by ~1e-4 relative):
  D_C(z) = c/H0 int_0^z dz' / E(z'),  E(z) = sqrt(Om0 (1+z)^3 + 1 - Om0),
  T_xy(z1, z2) = D_C(z2) - D_C(z1)           (transverse comoving, flat),
  d_xy(z1, z2) = T_xy(z1, z2) / (1 + z2)     (angular diameter distance).
Any object with H0 and Om0 attributes (an astropy FlatLambdaCDM, say) works.
"""

import numpy as np

C_KMS = 299792.458


class FlatLambdaCDM(object):
    def __init__(self, H0=67.66, Om0=0.30966):
        self.H0 = float(H0)
        self.Om0 = float(Om0)


class Background(object):
    """Distances in Mpc (Background.d_xy / T_xy, as used by MultiPlaneBulk)."""

    def __init__(self, cosmo=None, n_steps=8192):
        cosmo = FlatLambdaCDM() if cosmo is None else cosmo
        self._h0 = float(getattr(cosmo, "H0").value if hasattr(getattr(cosmo, "H0"), "value") else cosmo.H0)
        self._om = float(cosmo.Om0)
        self._n = int(n_steps)

    def _comoving(self, z):
        """D_C(z) by composite Simpson's rule (n_steps intervals)."""
        z = float(z)
        if z == 0.0:
            return 0.0
        zz = np.linspace(0.0, z, self._n + 1)
        f = 1.0 / np.sqrt(self._om * (1 + zz) ** 3 + 1 - self._om)
        h = z / self._n
        return C_KMS / self._h0 * h / 3.0 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())

    def T_xy(self, z_observer, z_source):
        return self._comoving(z_source) - self._comoving(z_observer)

    def d_xy(self, z_observer, z_source):
        z_obs = np.atleast_1d(np.asarray(z_observer, dtype=np.float64))
        z_src = np.atleast_1d(np.asarray(z_source, dtype=np.float64))
        out = np.array([(self._comoving(b) - self._comoving(a)) / (1 + b) for a, b in
                        np.broadcast(z_obs, z_src)])
        return out if np.ndim(z_observer) or np.ndim(z_source) else float(out[0])
