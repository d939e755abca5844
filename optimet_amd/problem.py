"""Cluster-solve problem description and the seeded synthetic workloads.

A ``Cluster`` carries what the reference's reader hands to its solver
(Reader.cpp:55-980 → Geometry / Excitation / ElectroMagnetic): sphere centres
and radii in metres, relative permittivities at ω and 2ω, the surface / bulk
second-harmonic susceptibilities, the background medium and the plane wave.

The Drude permittivity ε(ω) = 1 − ωp²/(ω² + iγω) is a material model the
reference does not have (its ElectroMagnetic.cpp:244-260 knows static, gold
and silicon models only); here it is evaluated per wavelength on the host and
handed over as constant ε(ω), ε(2ω) for that wavelength, exactly as a static
material would be.
"""
from dataclasses import dataclass, field, replace

import numpy as np

HBAR_EVS = 6.582119569e-16      # eV·s
C0 = 299792458.0                # constants.cpp:20
SI_KSIPPP = 65e-19              # ElectroMagnetic.cpp:238-240 (silicon SH constants)
SI_KSIPARPPAR = 3.5e-19
SI_GAMMA = 1.3e-19


@dataclass
class Cluster:
    xyz: np.ndarray                  # (P, 3) metres
    radius: np.ndarray               # (P,) metres
    eps_r: np.ndarray                # (P,) complex, at ω
    eps_r_sh: np.ndarray             # (P,) complex, at 2ω
    nmax: int
    wavelength: float                # metres
    mu_r: np.ndarray = None          # (P,) complex, default 1
    ksippp: np.ndarray = None        # (P,) complex
    ksiparppar: np.ndarray = None
    gamma: np.ndarray = None
    sh: bool = False
    bg_eps_r: complex = 1.0 + 0.0j
    bg_mu_r: complex = 1.0 + 0.0j
    theta_k: float = 0.0             # radians
    phi_k: float = 0.0
    etheta: complex = 1.0 + 0.0j
    ephi: complex = 0.0 + 0.0j
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        P = len(self.radius)
        self.xyz = np.asarray(self.xyz, dtype=np.float64).reshape(P, 3)
        self.radius = np.asarray(self.radius, dtype=np.float64)
        c = lambda v, d: np.broadcast_to(np.asarray(d if v is None else v, dtype=np.complex128), (P,)).copy()
        self.eps_r = c(self.eps_r, 1.0)
        self.eps_r_sh = c(self.eps_r_sh, 1.0)
        self.mu_r = c(self.mu_r, 1.0)
        self.ksippp = c(self.ksippp, 0.0)
        self.ksiparppar = c(self.ksiparppar, 0.0)
        self.gamma = c(self.gamma, 0.0)
        self.bg_eps_r = complex(self.bg_eps_r)
        self.bg_mu_r = complex(self.bg_mu_r)
        self.etheta = complex(self.etheta)
        self.ephi = complex(self.ephi)

    @property
    def nobj(self):
        return len(self.radius)

    @property
    def nf(self):
        """Harmonics per polarisation and object, nmax(nmax+2) (Tools.cpp:260)."""
        return self.nmax * (self.nmax + 2)

    @property
    def N(self):
        """Unknowns per frequency, 2·P·nf (Solver.h:132-135)."""
        return 2 * self.nf * self.nobj

    def hbm_bytes(self, fp64=True):
        """Dense matrix footprint, 16·N² (FP64) or 8·N² (complex FP32)."""
        return (16 if fp64 else 8) * self.N * self.N

    def with_wavelength(self, lam, drude=None):
        """Same cluster at another wavelength; re-evaluates Drude spheres."""
        c = replace(self, wavelength=float(lam), meta=dict(self.meta))
        c.__post_init__()
        if drude is None:
            drude = self.meta.get("drude")
        if drude is not None:
            mask = np.asarray(self.meta.get("drude_mask", np.ones(self.nobj, bool)))
            e1, e2 = drude_eps(lam, *drude)
            c.eps_r = np.where(mask, e1, c.eps_r)
            c.eps_r_sh = np.where(mask, e2, c.eps_r_sh)
        return c


def drude_eps(lam, wp_ev=9.0, gamma_ev=0.07):
    """Drude ε at ω = 2πc/λ and at 2ω; ħωp and ħγ in eV."""
    wp = wp_ev / HBAR_EVS
    g = gamma_ev / HBAR_EVS
    w = 2.0 * np.pi * C0 / lam
    eps = lambda om: 1.0 - wp * wp / (om * om + 1j * g * om)
    return complex(eps(w)), complex(eps(2.0 * w))


def random_cluster(P, rmin, rmax, cube, gap, seed, max_tries=2_000_000):
    """Seeded rejection sampling of non-overlapping spheres in a cube.

    Radii are uniform in [rmin, rmax]; centres are uniform in [0, cube]³ and a
    candidate is rejected if any surface gap is below ``gap``.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    radii = rng.uniform(rmin, rmax, size=P)
    xyz = np.zeros((P, 3))
    placed = 0
    tries = 0
    while placed < P:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("rejection sampling did not converge")
        c = rng.uniform(0.0, cube, size=3)
        if placed:
            d = np.linalg.norm(xyz[:placed] - c, axis=1)
            if np.any(d < radii[:placed] + radii[placed] + gap):
                continue
        xyz[placed] = c
        placed += 1
    return xyz, radii, rng


def _sh_draw(rng, P):
    f = rng.uniform(0.5, 1.5, size=(3, P))
    return f[0] * SI_KSIPPP, f[1] * SI_KSIPARPPAR, f[2] * SI_GAMMA


def config1():
    """3 spheres (a = 100 nm, ε = 12+0.1i) on the x axis at 0/300/600 nm, λ = 700 nm, nmax 6, FF only."""
    nm = 1e-9
    xyz = np.array([[0, 0, 0], [300, 0, 0], [600, 0, 0]], dtype=float) * nm
    return Cluster(xyz=xyz, radius=np.full(3, 100 * nm), eps_r=np.full(3, 12 + 0.1j),
                   eps_r_sh=np.full(3, 12 + 0.1j), nmax=6, wavelength=700 * nm, sh=False,
                   meta={"name": "config1-3spheres-nmax6"})


def drude_cluster(P, nmax, lam, rmin, rmax, cube, gap, seed, sh=True):
    nm = 1e-9
    xyz, radii, rng = random_cluster(P, rmin * nm, rmax * nm, cube * nm, gap * nm, seed)
    k3, k2, g = _sh_draw(rng, P)
    e1, e2 = drude_eps(lam)
    return Cluster(xyz=xyz, radius=radii, eps_r=np.full(P, e1), eps_r_sh=np.full(P, e2), nmax=nmax,
                   wavelength=lam, ksippp=k3, ksiparppar=k2, gamma=g, sh=sh,
                   meta={"drude": (9.0, 0.07), "seed": seed})


def config2():
    """64 Drude spheres, radii U[40,80] nm in a 1 µm cube (gap 10 nm), λ = 800 nm, nmax 8, ω + 2ω."""
    c = drude_cluster(64, 8, 800e-9, 40, 80, 1000, 10, seed=2000)
    c.meta["name"] = "config2-64spheres-nmax8-drude-ff+sh"
    return c


def config3_cluster():
    """16 Drude spheres (config-2 sampling), nmax 10; swept over 256 λ in [400, 900] nm."""
    c = drude_cluster(16, 10, 400e-9, 40, 80, 1000, 10, seed=3000)
    c.meta["name"] = "config3-16spheres-nmax10-sweep256"
    return c


def config3_wavelengths(n=256):
    return np.linspace(400e-9, 900e-9, n)


def config4():
    """512 Drude spheres, radii U[30,60] nm in a 3 µm cube (gap 10 nm), λ = 800 nm, nmax 6."""
    c = drude_cluster(512, 6, 800e-9, 30, 60, 3000, 10, seed=4000, sh=False)
    c.meta["name"] = "config4-512spheres-nmax6"
    return c


def serving_requests(n=2000, seed=5000, rate_hz=50.0, pmax=128, nmax_range=(4, 10), lam_max=64):
    """Poisson stream of cluster requests (seeded).

    Each request: P ~ U{2..pmax}, nmax ~ U{nmax_range}, 1..lam_max wavelengths in
    [400, 900] nm, Drude or constant-ε (12+0.1i) spheres, ω and 2ω.
    Returns a list of dicts with arrival time, cluster and wavelength list.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate_hz)
        P = int(rng.integers(2, pmax + 1))
        nmax = int(rng.integers(nmax_range[0], nmax_range[1] + 1))
        nl = int(rng.integers(1, lam_max + 1))
        lams = np.sort(rng.uniform(400e-9, 900e-9, size=nl))
        drude = bool(rng.integers(0, 2))
        cube = 1000.0 * max(1.0, (P / 64.0) ** (1.0 / 3.0))
        c = drude_cluster(P, nmax, float(lams[0]), 40, 80, cube, 10, seed=seed * 1000 + rid)
        if not drude:
            c.meta.pop("drude", None)
            c.eps_r = np.full(P, 12 + 0.1j)
            c.eps_r_sh = np.full(P, 12 + 0.1j)
        c.meta["name"] = f"request{rid}"
        out.append({"id": rid, "arrival": t, "cluster": c, "wavelengths": lams})
    return out
