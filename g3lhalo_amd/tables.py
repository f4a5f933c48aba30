"""Synthetic input tables for the NNMap / NMapMap integration path.

The reference reads nine precomputed tables (``src/NNMap_Model.h:51-60``):
five 1-D tables on the redshift grid and four 2-D tables stored z-major
(``hmf[z*N+m]``, ``P_lin[z*N+k]``, ``b_h[z*N+m]``, ``concentration[z*N+m]``).
Its ``doPrecomputations.x`` writes them; here they are generated directly on
the model grid with numpy.  This is input generation, not the integration
path:

* grids (``src/NNMap_Model.cu:253``, ``src/Function.cpp:89-105``,
  ``src/Function2D.cpp:66-95``):
  ``z_i = zmin + i (zmax-zmin)/N``, ``m_j = mmin exp(j ln(mmax/mmin)/N)``,
  ``k_j = kmin exp(j ln(kmax/kmin)/N)``, ``i, j = 0..N-1``;
* flat LCDM distances, ``dw/dz = c/H0/sqrt(E2)`` (``src/Cosmology.cpp:84-103``);
* lensing efficiency as the grid sum of ``src/Cosmology.cpp:131-151``;
* Eisenstein & Hu (1998) zero-baryon-wiggle transfer function, sigma8
  normalisation, linear growth of Carroll, Press & Turner (1992);
* Sheth-Tormen mass function and bias with A=0.322, q=0.707, p=0.3 in the
  form of ``src/HMF.cpp:7-49`` (mean density ``Omega_m(z) rho_crit``);
* concentration ``9/(1+z) (m/m*)^-0.13`` with ``m*`` the grid mass on a
  256-point grid over [1e10, 1e18] where ``|delta_c - sigma|`` is smallest
  (``src/Powerspectrum.cpp:117-150``).

Every table value is rounded to float32 precision (and stored as float64) so
the tables are reproducible bit for bit across machines whose libm differ in
the last places; :func:`checksum` fingerprints them.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

C_KMS = 2.99792e5          # src/constants.h:39
PI_REF = 3.14159           # src/constants.h:15
MPC_KM = 3.086e19          # src/Cosmology.h:28
G_MSUN = 1.327e11          # src/Cosmology.h:27
ARCMIN_IN_RAD = 2.909e-4   # src/constants.h:36


@dataclass
class Cosmo:
    Omega_m: float = 0.3
    Omega_Lambda: float = 0.7
    Omega_b: float = 0.045
    h: float = 0.7
    sigma8: float = 0.8
    ns: float = 0.96

    @property
    def H0(self) -> float:
        return 100.0 * self.h

    @property
    def rho_crit(self) -> float:
        # src/Cosmology.cpp:71, M_sun / Mpc^3
        return 3 * self.H0 * self.H0 * MPC_KM / 8 / PI_REF / G_MSUN


@dataclass
class Tables:
    nbins: int
    zmin: float
    zmax: float
    kmin: float
    kmax: float
    mmin: float
    mmax: float
    cosmo: Cosmo
    g: np.ndarray
    p_lens1: np.ndarray
    p_lens2: np.ndarray
    w: np.ndarray
    dwdz: np.ndarray
    hmf: np.ndarray
    P_lin: np.ndarray
    b_h: np.ndarray
    concentration: np.ndarray
    meta: dict = field(default_factory=dict)

    def arrays(self):
        return [self.g, self.p_lens1, self.p_lens2, self.w, self.dwdz,
                self.hmf, self.P_lin, self.b_h, self.concentration]


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).astype(np.float64))


def _e2(c: Cosmo, z):
    return c.Omega_m * (1 + z) ** 3 + c.Omega_Lambda


def _comoving(c: Cosmo, z: np.ndarray) -> np.ndarray:
    zf = np.linspace(0.0, float(np.max(z)) * 1.0001 + 1e-9, 200001)
    integrand = 1.0 / np.sqrt(_e2(c, zf))
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(zf))])
    return C_KMS / c.H0 * np.interp(z, zf, cum)


def _growth(c: Cosmo, z):
    e2 = _e2(c, z)
    om = c.Omega_m * (1 + z) ** 3 / e2
    ol = c.Omega_Lambda / e2
    d = 2.5 / (1 + z) * om / (om ** (4.0 / 7.0) - ol + (1 + 0.5 * om) * (1 + ol / 70.0))
    e0 = _e2(c, 0.0)
    om0 = c.Omega_m / e0
    ol0 = c.Omega_Lambda / e0
    d0 = 2.5 * om0 / (om0 ** (4.0 / 7.0) - ol0 + (1 + 0.5 * om0) * (1 + ol0 / 70.0))
    return d / d0


def _transfer_nowiggle(c: Cosmo, k: np.ndarray) -> np.ndarray:
    """Eisenstein & Hu (1998) eqs. 26-31, k in 1/Mpc."""
    h = c.h
    om = c.Omega_m * h * h
    ob = c.Omega_b * h * h
    fb = c.Omega_b / c.Omega_m
    theta = 2.7255 / 2.7
    s = 44.5 * np.log(9.83 / om) / np.sqrt(1 + 10 * ob ** 0.75)
    a_g = 1 - 0.328 * np.log(431 * om) * fb + 0.38 * np.log(22.3 * om) * fb * fb
    gamma = c.Omega_m * h * (a_g + (1 - a_g) / (1 + (0.43 * k * s) ** 4))
    q = k / h * theta * theta / gamma
    L0 = np.log(2 * np.e + 1.8 * q)
    C0 = 14.2 + 731.0 / (1 + 62.5 * q)
    return L0 / (L0 + C0 * q * q)


def _tophat(x):
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    small = x < 1e-3
    xs = x[small]
    out[small] = 1 - xs * xs / 10.0
    xl = x[~small]
    out[~small] = 3 * (np.sin(xl) - xl * np.cos(xl)) / xl ** 3
    return out


class _Linear:
    def __init__(self, c: Cosmo):
        self.c = c
        self.lnk = np.linspace(np.log(1e-6), np.log(1e4), 12001)
        self.k = np.exp(self.lnk)
        self.shape = self.k ** c.ns * _transfer_nowiggle(c, self.k) ** 2
        self.amp = 1.0
        s8 = self.sigma2_R0(8.0 / c.h)
        self.amp = c.sigma8 ** 2 / s8

    def p0(self, k):
        k = np.asarray(k, dtype=np.float64)
        return self.amp * k ** self.c.ns * _transfer_nowiggle(self.c, k) ** 2

    def sigma2_R0(self, R):
        R = np.atleast_1d(np.asarray(R, dtype=np.float64))
        w = _tophat(np.outer(R, self.k))
        f = self.amp * self.shape * self.k ** 3 * w * w / (2 * np.pi ** 2)
        dl = self.lnk[1] - self.lnk[0]
        res = (np.sum(f, axis=1) - 0.5 * (f[:, 0] + f[:, -1])) * dl
        return res if res.size > 1 else res[0]


def make_tables(nbins: int = 128, zmin: float = 0.001, zmax: float = 2.0, kmin: float = 1e-4,
                kmax: float = 1e4, mmin: float = 1e10, mmax: float = 1e17,
                cosmo: Cosmo | None = None, lens1=(0.4, 0.1), lens2=(0.4, 0.1),
                src_z0: float = 0.5) -> Tables:
    """Tables on the model grid (see module docstring)."""
    c = cosmo or Cosmo()
    N = int(nbins)
    i = np.arange(N, dtype=np.float64)
    z = zmin + i * (zmax - zmin) / N
    m = np.exp(np.log(mmin) + i * np.log(mmax / mmin) / N)
    k = np.exp(np.log(kmin) + i * np.log(kmax / kmin) / N)
    dz = (zmax - zmin) / N

    w = _comoving(c, z)
    dwdz = C_KMS / c.H0 / np.sqrt(_e2(c, z))

    def gauss(mu, sig):
        p = np.exp(-0.5 * ((z - mu) / sig) ** 2)
        return p / (p.sum() * dz)

    p1 = gauss(*lens1)
    p2 = gauss(*lens2)
    ns = z * z * np.exp(-(z / src_z0) ** 1.5)
    ns = ns / (ns.sum() * dz)
    g = np.zeros(N)
    for a in range(N):
        sel = np.arange(a, N)
        g[a] = np.sum(ns[sel] * (w[sel] - w[a]) / w[sel]) * dz

    lin = _Linear(c)
    D = _growth(c, z)
    P = (D[:, None] ** 2) * lin.p0(k)[None, :]

    e2 = _e2(c, z)
    om_z = c.Omega_m * (1 + z) ** 3 / e2
    rho_mean0 = c.Omega_m * c.rho_crit
    rho_z = om_z * c.rho_crit
    delta_c = 1.686 * (1 + 0.0123 * np.log10(om_z))

    def sigma2_of_m(mm):
        R = (3 * mm / (4 * np.pi * rho_mean0)) ** (1.0 / 3.0)
        return lin.sigma2_R0(R)

    s2_0 = sigma2_of_m(m)
    eps = 1e-3
    ds2dlnm = (sigma2_of_m(m * np.exp(eps)) - sigma2_of_m(m * np.exp(-eps))) / (2 * eps)
    A, q, p = 0.322, 0.707, 0.3
    s2 = (D[:, None] ** 2) * s2_0[None, :]
    ds2dm = (D[:, None] ** 2) * (ds2dlnm / m)[None, :]
    nu2 = delta_c[:, None] ** 2 / s2
    hmf = (-rho_z[:, None] / m[None, :] / s2 * ds2dm * A * (1 + (q * nu2) ** (-p))
           * np.sqrt(q * nu2 / 2 / PI_REF) * np.exp(-0.5 * q * nu2))
    bh = 1 + 1.0 / delta_c[:, None] * (q * nu2 - 1 + 2 * p / (1 + (q * nu2) ** p))

    mgrid = np.exp(np.log(1e10) + np.arange(256) * np.log(1e18 / 1e10) / 256)
    s_star = np.sqrt(sigma2_of_m(mgrid))
    conc = np.empty((N, N))
    for a in range(N):
        sig = D[a] * s_star
        mstar = mgrid[int(np.argmin(np.abs(delta_c[a] - sig)))]
        conc[a] = 9.0 / (1 + z[a]) * (m / mstar) ** (-0.13)

    t = Tables(N, zmin, zmax, kmin, kmax, mmin, mmax, c,
               _f32(g), _f32(p1), _f32(p2), _f32(w), _f32(dwdz),
               _f32(hmf.reshape(-1)), _f32(P.reshape(-1)), _f32(bh.reshape(-1)),
               _f32(conc.reshape(-1)))
    t.meta = {"z": z, "m": m, "k": k}
    return t


def checksum(t: Tables) -> str:
    hsh = hashlib.sha256()
    for a in t.arrays():
        hsh.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return hsh.hexdigest()


# HOD vectors are {f, alpha, M_th, sigma, M', beta} (src/HOD.cu:15-20).
# The fiducial set: f=0.3, M_th=1e12, sigma=0.5, M'=10^13.3, satellite slope
# beta=1, central amplitude alpha=1.
FIDUCIAL_HOD = (0.3, 1.0, 1e12, 0.5, 10 ** 13.3, 1.0)


def hod_prior_draws(n: int, seed: int = 20220405) -> np.ndarray:
    """Seeded uniform prior draws: f in [0.1,0.9], log10 M_th in [11,13],
    sigma in [0.1,1], log10 M' in [12,14.5], slope beta in [0.5,1.5];
    central amplitude fixed at 1."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.1, 0.9, n)
    mth = 10 ** rng.uniform(11, 13, n)
    sig = rng.uniform(0.1, 1.0, n)
    mp = 10 ** rng.uniform(12, 14.5, n)
    beta = rng.uniform(0.5, 1.5, n)
    return np.stack([f, np.ones(n), mth, sig, mp, beta], axis=1)


def aperture_grid() -> np.ndarray:
    """Ten log-spaced apertures 0.5 .. 32 arcmin."""
    return 0.5 * 64.0 ** (np.arange(10) / 9.0)


def triplets_550() -> np.ndarray:
    """All (theta1 <= theta2) pairs x 10 theta3 on the aperture grid, arcmin."""
    a = aperture_grid()
    out = [(a[i], a[j], a[k]) for i in range(10) for j in range(i, 10) for k in range(10)]
    return np.asarray(out, dtype=np.float64)


def triplets_55() -> np.ndarray:
    """(a_i, a_j, a_j) for i <= j: 55 triplets, 10 of them equal-aperture."""
    a = aperture_grid()
    return np.asarray([(a[i], a[j], a[j]) for i in range(10) for j in range(i, 10)],
                      dtype=np.float64)


def triplets_equal_4() -> np.ndarray:
    return np.asarray([(1, 1, 1), (2, 2, 2), (4, 4, 4), (8, 8, 8)], dtype=np.float64)


def write_reference_files(t: Tables, directory: str, hods=(FIDUCIAL_HOD,), thetas=None) -> dict:
    """Writes the tables in the reference's ASCII input formats: 1-D files of
    "z value" rows on the model grid plus one row at zmax (the reference's
    Function::at returns its default at the last abscissa, so the grid's last
    point needs a successor), 2-D files of "z m value" / "z k value" rows in
    z-major order, a cosmology file, HOD files and a thetas file.  Values are
    written with 17 significant digits so they read back bit for bit."""
    import os
    os.makedirs(directory, exist_ok=True)
    z, m, k = t.meta["z"], t.meta["m"], t.meta["k"]
    n = t.nbins
    zx = np.append(t.zmin + np.arange(n, dtype=np.float64) * (t.zmax - t.zmin) / n, t.zmax)
    files = {}

    def one_d(name, y):
        p = os.path.join(directory, name)
        with open(p, "w") as f:
            for xi, yi in zip(zx, np.append(y, y[-1])):
                f.write(f"{xi:.17g} {yi:.17g}\n")
        files[name] = p

    def two_d(name, x2, y):
        p = os.path.join(directory, name)
        yy = y.reshape(n, n)
        with open(p, "w") as f:
            for i in range(n):
                for j in range(n):
                    f.write(f"{z[i]:.17g} {x2[j]:.17g} {yy[i, j]:.17g}\n")
        files[name] = p

    one_d("g.dat", t.g)
    one_d("pz_lens1.dat", t.p_lens1)
    one_d("pz_lens2.dat", t.p_lens2)
    one_d("w.dat", t.w)
    one_d("dwdz.dat", t.dwdz)
    two_d("hmf.dat", m, t.hmf)
    two_d("Plin.dat", k, t.P_lin)
    two_d("b_h.dat", m, t.b_h)
    two_d("concentration.dat", m, t.concentration)
    c = t.cosmo
    p = os.path.join(directory, "cosmo.param")
    with open(p, "w") as f:
        f.write(f"Omega_m {c.Omega_m!r}\nOmega_Lambda {c.Omega_Lambda!r}\nOmega_b {c.Omega_b!r}\n"
                f"sigma8 {c.sigma8!r}\nH0 {c.H0!r}\nns {c.ns!r}\n")
    files["cosmo.param"] = p
    for i, h in enumerate(hods):
        p = os.path.join(directory, f"hod{i + 1}.param")
        f_, alpha, mth, sig, mp, beta = (float(v) for v in h)
        with open(p, "w") as f:
            f.write(f"M_min {mth!r}\nsigma {sig!r}\nM' {mp!r}\nbeta {beta!r}\nalpha {alpha!r}\n"
                    f"f {f_!r}\n")
        files[f"hod{i + 1}.param"] = p
    if thetas is not None:
        p = os.path.join(directory, "thetas.dat")
        with open(p, "w") as f:
            for row in thetas:
                f.write(" ".join(f"{v:.17g}" for v in row) + "\n")
        files["thetas.dat"] = p
    return files
