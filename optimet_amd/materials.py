"""Wavelength-dependent sphere materials of OPTIMET's input format.

ElectroMagnetic::update(lambda) (ElectroMagnetic.cpp:244-260) re-evaluates a
sphere's permittivities at omega and 2 omega, and for the gold model its
second-harmonic susceptibilities, every time the wavelength changes.  The
reader (Reader.cpp:600-666) knows three epsilon types:

* "relative": constant values (nothing to re-evaluate);
* "SiliconModel": the refractive index n and extinction coefficient k of
  crystalline silicon tabulated from 0.25 um to 1.45 um in 0.01 um steps (the
  Schinke et al. data the reference names), linearly interpolated at lambda
  and at lambda / 2, eps = (n^2 - k^2) + 2 i n k, and the constant surface /
  bulk susceptibilities 65e-19, 3.5e-19, 1.3e-19
  (ElectroMagnetic.cpp:144-242);
* "GoldModel": a five-term rational fit of gold's permittivity in omega,
  eps = 1 + sum_i (a0_i - i omega a1_i) / (b0_i - i omega b1_i - omega^2 b2_i),
  at omega and 2 omega, and the hydrodynamic susceptibilities
  chi_perp-perp-perp = -(a/4) (eps - 1) e / (m_e omega^2),
  chi_par-perp-par = -(b/2) (...), gamma = -(d/8) (...)
  (ElectroMagnetic.cpp:69-139), with a, b, d from the case file.

Both models are evaluated on the host, once per wavelength, in the same
double-precision operation order as the reference; the wavelength they see is
2 pi / (2 pi / lambda), as Excitation::lambda() recovers it from the wave
number (Excitation.h:99).  A silicon wavelength outside the table (lambda or
lambda / 2 beyond [0.25, 1.45] um) reads past the reference's arrays; here it
raises.
"""
import math

import numpy as np

PI = 3.14159265358979323846      # constants.cpp:19
C0 = 299792458.0                 # constants.cpp:20
M_E = 9.10938356e-31             # ElectroMagnetic.cpp:133
Q_E = 1.602176e-19               # ElectroMagnetic.cpp:135

# Crystalline silicon, n and k at 0.25 + 0.01 i um, i = 0..120.
SI_N = np.array([
    1.6370, 1.7370, 2.0300, 2.8400, 4.1850, 5.0490, 5.0910, 5.0850, 5.1350, 5.2450,
    5.4230, 5.9140, 6.8200, 6.5870, 6.0250, 5.6230, 5.3410, 5.1100, 4.9320, 4.7900,
    4.6730, 4.5720, 4.4850, 4.4120, 4.3490, 4.2890, 4.2350, 4.1870, 4.1450, 4.1030,
    4.0730, 4.0380, 4.0060, 3.9770, 3.9540, 3.9310, 3.9080, 3.8880, 3.8690, 3.8510,
    3.8350, 3.8170, 3.8050, 3.7910, 3.7760, 3.7650, 3.7530, 3.7410, 3.7300, 3.7190,
    3.7120, 3.7010, 3.6930, 3.6840, 3.6770, 3.6690, 3.6620, 3.6550, 3.6460, 3.6410,
    3.6360, 3.6280, 3.6220, 3.6170, 3.6130, 3.6100, 3.6040, 3.5980, 3.5970, 3.5900,
    3.5840, 3.5840, 3.5780, 3.5820, 3.5790, 3.5750, 3.5720, 3.5680, 3.5650, 3.5620,
    3.5590, 3.5560, 3.5530, 3.5490, 3.5470, 3.5450, 3.5420, 3.5400, 3.5370, 3.5340,
    3.5330, 3.5300, 3.5270, 3.5260, 3.5240, 3.5220, 3.5200, 3.5180, 3.5170, 3.5150,
    3.5130, 3.5120, 3.5090, 3.5090, 3.5060, 3.5050, 3.5030, 3.5020, 3.5010, 3.5000,
    3.4990, 3.4970, 3.4960, 3.4960, 3.4960, 3.4930, 3.4920, 3.4920, 3.4900, 3.4880,
    3.4870,
])
SI_K = np.array([
    3.5889, 3.9932, 4.5958, 5.1961, 5.3124, 4.2900, 3.6239, 3.2824, 3.0935, 2.9573,
    2.9078, 2.9135, 2.1403, 0.9840, 0.5031, 0.3263, 0.2413, 0.1769, 0.1377, 0.1120,
    0.0954, 0.0791, 0.0702, 0.0598, 0.0538, 0.0485, 0.0438, 0.0395, 0.0348, 0.0299,
    0.0280, 0.0266, 0.0237, 0.0219, 0.0201, 0.0185, 0.0173, 0.0168, 0.0163, 0.0147,
    0.0144, 0.0136, 0.0128, 0.0120, 0.0113, 0.0106, 0.0100, 0.0093, 0.0087, 0.0082,
    0.0076, 0.0071, 0.0066, 0.0061, 0.0057, 0.0053, 0.0049, 0.0045, 0.0041, 0.0038,
    0.0035, 0.0032, 0.0029, 0.0026, 0.0023, 0.0021, 0.0019, 0.0017, 0.0015, 0.0013,
    0.0011, 9.8243e-04, 8.4060e-04, 7.1334e-04, 5.9638e-04, 4.9020e-04, 3.9616e-04, 3.1437e-04, 2.4048e-04, 1.7959e-04,
    1.3043e-04, 9.2450e-05, 6.7820e-05, 5.2168e-05, 3.9770e-05, 3.0217e-05, 2.2913e-05, 1.7068e-05, 1.2382e-05, 8.6210e-06,
    5.6876e-06, 3.4275e-06, 1.7653e-06, 5.5561e-07, 2.3153e-07, 1.3904e-07, 8.0863e-08, 4.7940e-08, 2.7132e-08, 1.4318e-08,
    5.8798e-09, 2.3352e-09, 1.2714e-09, 7.5284e-10, 4.4799e-10, 2.7228e-10, 1.5856e-10, 8.7196e-11, 4.2039e-11, 1.8128e-11,
    1.0428e-11, 6.2911e-12, 3.9030e-12, 2.6367e-12, 1.7377e-12, 1.0428e-12, 6.0422e-13, 4.2895e-13, 2.0381e-13, 1.3785e-13,
    1.0901e-13,
])
# Five-term rational fit of gold's permittivity (SI units of omega).
AU_A0 = (2.000003399882560, 1.782388034422510e+32, 9.571140818411450e+26, 3.141025290600320e+24,
         5.056282927859510e+31)
AU_A1 = (0.0, 0.0, 8.034165109695690e+15, 1.060027902520820e+14, 2.176317566053640e+16)
AU_B0 = (1.0, 0.0, 1.398566311205070e+26, 5.984581206741880e+23, 1.707510287416960e+31)
AU_B1 = (1.326291192399820e-15, 1.122727361975370e+14, 7.280057361739550e+15, 4.393809682455200e+15,
         3.256258123271410e+15)
AU_B2 = (0.0, 1.0, 1.0, 1.0, 1.0)


def model_wavelength(lam):
    """The wavelength the material models see: 2 pi / (2 pi / lambda)."""
    return 2 * PI / (2 * PI / lam)


def _si_interp(lum):
    for b in range(SI_K.size):
        s1 = 0.25 + b * 0.01
        s2 = 0.25 + (b + 1) * 0.01
        if s1 <= lum <= s2:
            break
    else:
        b = SI_K.size
    if b + 1 >= SI_K.size:
        raise ValueError(f"SiliconModel: {lum} um lies outside the tabulated 0.25-1.45 um")
    n = SI_N[b] + ((SI_N[b + 1] - SI_N[b]) / (s2 - s1)) * (lum - s1)
    k = SI_K[b] + ((SI_K[b + 1] - SI_K[b]) / (s2 - s1)) * (lum - s1)
    return float(n), float(k)


def silicon(lam):
    """(eps, eps_sh, ksippp, ksiparppar, gamma) of the silicon model at vacuum
    wavelength lam (m)."""
    lff = model_wavelength(lam) * 1e6
    nf, kf = _si_interp(lff)
    ns, ks = _si_interp(lff / 2.0)
    eps = complex(nf ** 2 - kf ** 2, 2.0 * nf * kf)
    eps_sh = complex(ns ** 2 - ks ** 2, 2.0 * ns * ks)
    return eps, eps_sh, 65e-19 + 0j, 3.5e-19 + 0j, 1.3e-19 + 0j


def gold(lam, a, b, d):
    """(eps, eps_sh, ksippp, ksiparppar, gamma) of the gold model at vacuum
    wavelength lam (m), hydrodynamic parameters a, b, d (complex)."""
    f = C0 / model_wavelength(lam)
    w = 2 * PI * f
    s_ff, s_sh = 0j, 0j
    for i in range(5):
        z1 = w * complex(0.0, -1.0)
        z2 = 2.0 * w * complex(0.0, -1.0)
        s_ff += (AU_A0[i] + z1 * AU_A1[i]) / (AU_B0[i] + z1 * AU_B1[i] + (z1 * z1) * AU_B2[i])
        s_sh += (AU_A0[i] + z2 * AU_A1[i]) / (AU_B0[i] + z2 * AU_B1[i] + (z2 * z2) * AU_B2[i])
    eps = 1.0 + s_ff
    eps_sh = 1.0 + s_sh
    scale = (eps - 1.0) * Q_E / (M_E * math.pow(2.0 * PI * f, 2.0))
    return eps, eps_sh, -(complex(a) / 4.0) * scale, -(complex(b) / 2.0) * scale, -(complex(d) / 8.0) * scale


class Material:
    """A sphere's material: kind "relative" (constant), "SiliconModel",
    "GoldModel" (params a, b, d) or "Drude" (params wp, gamma in eV)."""

    def __init__(self, kind, eps=1.0, eps_sh=0.0, ksippp=0.0, ksiparppar=0.0, gamma=0.0, params=None):
        self.kind = kind
        self.const = (complex(eps), complex(eps_sh), complex(ksippp), complex(ksiparppar), complex(gamma))
        self.params = params or ()

    @property
    def dispersive(self):
        return self.kind != "relative"

    @property
    def varying_susceptibilities(self):
        return self.kind == "GoldModel"

    def at(self, lam):
        """(eps, eps_sh, ksippp, ksiparppar, gamma) at vacuum wavelength lam."""
        if self.kind == "relative":
            return self.const
        if self.kind == "SiliconModel":
            return silicon(lam)
        if self.kind == "GoldModel":
            return gold(lam, *self.params)
        if self.kind == "Drude":
            from .problem import drude_eps

            e1, e2 = drude_eps(lam, *self.params)
            return (e1, e2) + self.const[2:]
        raise ValueError(f"unknown material kind {self.kind!r}")
