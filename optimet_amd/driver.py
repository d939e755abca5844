"""Case driver: OPTIMET's XML input and its outputs, over the MI355X backend.

``python -m optimet_amd.driver case.xml [--kind lu32ir|lu64|gmres64|aca]``
does what ``Optimet3D case`` does for a sphere cluster (main.cpp ->
Simulation::run, Simulation.cpp:33-63).  ``<ACA compression="yes"/>`` selects
the ACA-compressed operator with the reference's GMRES settings
(PreconditionedMatrixSolver.h:47-56), as the reference's solve() does; without
it the mixed-precision LU (or the --kind given) solves the dense systems.

* ``<output type="response">`` with ``<scan><wavelength initial= final=
  stepsize=/></scan>``: the wavelength scan of Simulation::scan_wavelengths
  (Simulation.cpp:604-685), solved as batched sweeps on the device, written as
  ``<case>_AbsorptionCS_FF.dat`` (C_ext - C_sca, Simulation.cpp:659),
  ``<case>_ScatteringCS_FF.dat`` and, with SH sources,
  ``<case>_ScatteringCS_SH.dat`` and ``<case>_AbsorptionCS_SH.dat``: one line
  per wavelength, "lambda<TAB>value", lambda in metres, six significant digits
  (the default formatting of the reference's std::ofstream);
* ``<output type="field">`` with ``<grid><x min= max= steps=/>...</grid>``: the
  field maps of Simulation::field_simulation (Simulation.cpp:336-361) on the
  regular grid of OutputGrid (OutputGrid.cpp:132-159, point (i, j, k) =
  (x_min + i dx, ...) + 1e-12 m, x fastest), written to ``<case>_FF.npz`` and
  ``<case>_SH.npz`` with the datasets of the reference's HDF5 files
  (OutputGrid.cpp:57-88: Field_E/X/real, .../imag, Y, Z, ABS/abs, and
  Field_H/...), each an (n_x, n_y, n_z) array.  HDF5 itself is not written:
  this image has no HDF5 Python binding.

The reader follows Reader.cpp: nmax from ``<simulation><harmonics nmax=>``
(nmaxS = nmax, :66-67); spheres with ``<cartesian x= y= z=/>`` or ``<spherical
rrr= the= phi=/>`` (nm, rad), ``<properties radius=>`` (nm), ``<mu
type="relative">`` and an ``<epsilon>`` of type "relative" (with
``epsilon_SH``, ``ksippp``, ``ksiparppar``, ``gamma``), "SiliconModel" or
"GoldModel" (:590-666, the models in materials.py); the lattices of
``<structure type="cube"|"surface"|"GaAssurf"|"GaAscube">`` (:103-540); the
plane wave (:782-831: wavelength in nm, propagation angles in degrees,
E_theta, E_phi, ``<SHsources condition="yes"/>``).  Its quirks are kept: the
background values are read only when the background's ``type`` is not
"relative" (:84, :107); a lattice lists its spheres in the order its
builder's push loop leaves them, the lattice point at the origin last (:166-175,
:298-311, :520-536); the scan has int(final - initial) / stepsize + 1
wavelengths (:886-891); each sphere is checked for overlap against the spheres
already pushed when it is pushed (Geometry.cpp:39-51).  One quirk is refused
instead of reproduced: a scan of a single wavelength makes the reference
divide 0 by 0 and solve at lambda = NaN (Reader.cpp:890-891 with
Simulation.cpp:641-644); here it is an error.  Arbitrary shapes are outside
this backend and raise.  An ``epsilon type="Drude"`` element (``wp`` and
``gamma`` in eV) is this build's addition.
"""
import argparse
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np

from .materials import Material
from .problem import Cluster

NM = 1e-9


class CaseError(ValueError):
    pass


def _c(node, name):
    """Complex value.real / value.imag attributes of child `name`; 0 where
    absent, as pugixml's as_double() of a missing attribute."""
    ch = node.find(name) if node is not None else None
    if ch is None:
        return 0j
    return complex(float(ch.get("value.real", 0.0)), float(ch.get("value.imag", 0.0)))


def _f(node, attr):
    if node is None:
        return 0.0
    v = node.get(attr)
    return float(v) if v not in (None, "") else 0.0


def _cart(r, the, phi):
    return [r * math.sin(the) * math.cos(phi), r * math.sin(the) * math.sin(phi), r * math.cos(the)]


class Sphere:
    def __init__(self, xyz, radius, material, mu):
        self.xyz, self.radius, self.material, self.mu = list(xyz), float(radius), material, complex(mu)

    def copy(self):
        return Sphere(self.xyz, self.radius, self.material, self.mu)


def read_scatterer(node, i):
    """read_scatterer (Reader.cpp:590-666) for a sphere."""
    if node is None or node.get("type") != "sphere":
        raise CaseError(f"object {i}: only spheres are supported by this backend")
    cart, sph = node.find("cartesian"), node.find("spherical")
    if cart is not None:
        xyz = [_f(cart, "x") * NM, _f(cart, "y") * NM, _f(cart, "z") * NM]
    elif sph is not None:
        xyz = _cart(_f(sph, "rrr") * NM, _f(sph, "the"), _f(sph, "phi"))
    else:
        xyz = [0.0, 0.0, 0.0]
    radius = _f(node.find("properties"), "radius") * NM
    en, mun = node.find("epsilon"), node.find("mu")
    if en is None and mun is None:
        raise CaseError(f"object {i}: no material (epsilon / mu) given")
    if mun is None or mun.get("type") != "relative":
        raise CaseError('The type for mu must be "relative"')
    mu = _c(node, "mu")
    kind = en.get("type") if en is not None else None
    if kind == "relative":
        mat = Material("relative", _c(node, "epsilon"), _c(node, "epsilon_SH"), _c(node, "ksippp"),
                       _c(node, "ksiparppar"), _c(node, "gamma"))
    elif kind == "SiliconModel":
        mat = Material("SiliconModel")
    elif kind == "GoldModel":
        p = en.find("parameters")
        cp = lambda n: complex(_f(p, n + ".real"), _f(p, n + ".imag"))
        mat = Material("GoldModel", params=(cp("a"), cp("b"), cp("d")))
    elif kind == "Drude":
        mat = Material("Drude", ksippp=_c(node, "ksippp"), ksiparppar=_c(node, "ksiparppar"),
                       gamma=_c(node, "gamma"), params=(_f(en, "wp"), _f(en, "gamma")))
    else:
        raise CaseError("Unknown type for epsilon")
    if not radius > 0:
        raise CaseError(f"object {i}: the sphere needs a positive radius")
    return Sphere(xyz, radius, mat, mu)


def _push(spheres, s):
    """Geometry::pushObject (Geometry.cpp:39-51): overlap against every sphere
    already pushed, then append."""
    for o in spheres:
        if math.dist(o.xyz, s.xyz) <= s.radius + o.radius:
            raise CaseError(f"The sphere at ({s.xyz[0]}, {s.xyz[1]}, {s.xyz[2]}) overlaps with the one at "
                            f"({o.xyz[0]}, {o.xyz[1]}, {o.xyz[2]}), with radii {s.radius} and {o.radius}")
    spheres.append(s.copy())


def _lattice(proto, pts, radii=None, skip_origin=False, first=1):
    """The builders' push loop: the prototype is pushed, then for each lattice
    point the last sphere is moved there (and given its radius) and a fresh
    prototype pushed, so the prototype's own position ends up last."""
    spheres = []
    _push(spheres, proto)
    for i in range(first, len(pts)):
        if skip_origin and pts[i] == (0.0, 0.0, 0.0):
            continue
        if radii is not None:
            spheres[-1].radius = radii[i]
        spheres[-1].xyz = list(pts[i])
        _push(spheres, proto)
    return spheres


def _zincblende_cell(a):
    X1 = [a, a, a / 2.0, a / 2.0, 0.0, 0.0, a / 2.0, a, a / 2.0, 0.0, 0.0, a, a, 0.0, 3.0 * a / 4.0, a / 4.0,
          3.0 * a / 4.0, a / 4.0]
    Y1 = [0.0, a / 2.0, a / 2.0, 0.0, a / 2.0, a, a, a, a / 2.0, 0.0, 0.0, 0.0, a, a, a / 4.0, 3.0 * a / 4.0,
          3.0 * a / 4.0, a / 4.0]
    Z1 = [0.0, a / 2.0, 0.0, a / 2.0, a / 2.0, 0.0, a / 2.0, a, a, a, 0.0, a, 0.0, a, a / 4.0, a / 4.0,
          3.0 * a / 4.0, 3.0 * a / 4.0]
    return X1, Y1, Z1


def read_structure(sn):
    """read_structure (Reader.cpp:103-540): cube, surface and the two
    zinc-blende (GaAs) lattices."""
    kind = sn.get("type")
    proto = read_scatterer(sn.find("object"), 0)
    props = sn.find("properties")
    if kind in ("cube", "surface"):
        No = int(_f(props, "points"))
        d = _f(props, "distance") * NM
        pts = []
        for k in range(No if kind == "cube" else 1):
            for j in range(No):
                for i in range(No):
                    pts.append((d * float(i), d * float(j), d * float(k)))
        return _lattice(proto, pts)
    if kind in ("GaAssurf", "GaAscube"):
        No = int(_f(props, "No_cells"))
        a = _f(props, "cell_size") * NM
        R = _f(sn.find("object/properties"), "radius") * NM
        ratio = 2.0 if kind == "GaAssurf" else 1.7
        X1, Y1, Z1 = _zincblende_cell(a)
        pts, radii = [], []
        for t in range(No if kind == "GaAscube" else 1):
            for i in range(No):
                for j in range(No):
                    for k in range(18):
                        keep = ((i == 0 or X1[k] > 0.0) and (j == 0 or Y1[k] > 0.0) and (t == 0 or Z1[k] > 0.0))
                        if keep:
                            pts.append((X1[k] + a * float(i), Y1[k] + a * float(j), Z1[k] + a * float(t)))
                            radii.append(R if k < 14 else R / ratio)
        return _lattice(proto, pts, radii, skip_origin=True, first=0)
    raise CaseError(f"structure type {kind!r} is not supported")


class Case:
    """A parsed case file: the spheres, background, plane wave and output
    request (kind "response" with wavelengths, or "field" with a grid)."""

    def __init__(self, name, nmax, spheres, bg, wave, output, wavelengths=None, grid=None, aca=False,
                 projection=False):
        self.name, self.nmax, self.spheres, self.bg, self.wave = name, nmax, spheres, bg, wave
        self.output, self.wavelengths, self.grid = output, wavelengths, grid
        self.aca, self.projection = aca, projection

    @property
    def sh(self):
        return self.wave["sh"]

    def at(self, lam):
        """The cluster at vacuum wavelength lam, every material re-evaluated
        (ElectroMagnetic::update, ElectroMagnetic.cpp:244-260)."""
        m = [s.material.at(lam) for s in self.spheres]
        col = lambda i: np.array([v[i] for v in m])
        w = self.wave
        return Cluster(xyz=np.array([s.xyz for s in self.spheres]), radius=np.array([s.radius for s in self.spheres]),
                       eps_r=col(0), eps_r_sh=col(1), nmax=self.nmax, wavelength=float(lam),
                       mu_r=np.array([s.mu for s in self.spheres]), ksippp=col(2), ksiparppar=col(3), gamma=col(4),
                       sh=w["sh"], bg_eps_r=self.bg[0], bg_mu_r=self.bg[1], theta_k=w["theta"], phi_k=w["phi"],
                       etheta=w["etheta"], ephi=w["ephi"], meta={"name": self.name})

    @property
    def cluster(self):
        """The cluster at the source wavelength."""
        return self.at(self.wave["lambda"])


def _parse(path):
    """pugixml accepts several top-level elements; ElementTree wants one."""
    text = open(path).read()
    text = re.sub(r"<\?xml[^>]*\?>", "", text, count=1)
    return ET.fromstring("<case>" + text + "</case>")


def read_case(path):
    root = _parse(path)
    sim = root.find("simulation")
    if sim is None:
        raise CaseError("Simulation parameters not defined!")
    nmax = int(_f(sim.find("harmonics"), "nmax"))
    if nmax < 1:
        raise CaseError("harmonics nmax must be at least 1")
    aca = sim.find("ACA") is not None and sim.find("ACA").get("compression") == "yes"
    geo = root.find("geometry")
    if geo is None:
        raise CaseError("Geometry not defined!")
    if geo.find("structure") is not None:
        spheres = read_structure(geo.find("structure"))
    else:
        spheres = []
        for i, node in enumerate(geo.findall("object")):
            _push(spheres, read_scatterer(node, i))
    bg = (1.0 + 0j, 1.0 + 0j)
    bgn = geo.find("background")
    if bgn is not None and bgn.get("type", "") != "relative":  # Reader.cpp:84 (strcmp != 0)
        bg = (_c(bgn, "epsilon"), _c(bgn, "mu"))
    if not spheres:
        raise CaseError("No scatterers defined in input")
    src = root.find("source")
    if src is None:
        raise CaseError("Source not defined!")
    lam = _f(src.find("wavelength"), "value") * NM
    if not lam > 0:
        raise CaseError("the source needs a positive wavelength")
    prop, pol = src.find("propagation"), src.find("polarization")
    shn = src.find("SHsources")
    wave = {"lambda": lam, "theta": _f(prop, "theta") * math.pi / 180.0, "phi": _f(prop, "phi") * math.pi / 180.0,
            "etheta": complex(_f(pol, "Etheta.real"), _f(pol, "Etheta.imag")),
            "ephi": complex(_f(pol, "Ephi.real"), _f(pol, "Ephi.imag")),
            "sh": shn is not None and shn.get("condition") == "yes"}
    name = os.path.splitext(os.path.basename(path))[0]
    out = root.find("output")
    if out is None:
        raise CaseError("Output not defined!")
    otype = out.get("type")
    if otype == "response":
        if out.find("scan/radius") is not None:
            # outputType 12 / 112 (Reader.cpp:893-902): Simulation::run has no
            # branch for them and prints "Nothing to do?" (Simulation.cpp:59-62)
            raise CaseError("radius scans are not run by the reference (Simulation::run has no branch for them)")
        wl = out.find("scan/wavelength")
        if wl is None:
            raise CaseError("a response output needs <scan><wavelength .../></scan>")
        lam_i, lam_f = _f(wl, "initial"), _f(wl, "final")
        stepsize = int(_f(wl, "stepsize"))
        if stepsize == 0:
            raise CaseError("scan stepsize must be non-zero")
        span = int(lam_f - lam_i)
        steps = int(span / stepsize) + 1  # Reader.cpp:886-891 (C integer division truncates toward 0)
        if steps < 2:
            raise CaseError("a scan of a single wavelength divides 0 by 0 in the reference (lambda = NaN); "
                            "give at least two wavelengths")
        lami, lamf = lam_i * NM, lam_f * NM
        d = (lamf - lami) / (steps - 1)
        lams = np.array([lami + i * d for i in range(steps)])
        return Case(name, nmax, spheres, bg, wave, "response", wavelengths=lams, aca=aca)
    if otype == "field":
        g = out.find("grid")
        grid = []
        for ax in ("x", "y", "z"):
            a = g.find(ax) if g is not None else None
            n = int(_f(a, "steps"))
            if n < 2:
                raise CaseError("At least 2 output steps must be specified for any given axis")
            grid.append((_f(a, "min") * NM, _f(a, "max") * NM, n))
        proj = out.find("projection")
        projection = proj is not None and proj.get("spherical") == "true"  # Reader.cpp:860-861
        return Case(name, nmax, spheres, bg, wave, "field", grid=grid, aca=aca, projection=projection)
    raise CaseError(f"output type {otype!r} is not supported by this backend")


def grid_points(grid):
    """OutputGrid's points (OutputGrid.cpp:132-159): iterator i -> cursor
    (i mod nx, (i / nx) mod ny, i / (nx ny)), coordinate min + cursor step +
    1e-12; returns (n, 3) in iterator order and the (nx, ny, nz) shape."""
    (x0, x1, nx), (y0, y1, ny), (z0, z1, nz) = grid
    dx, dy, dz = abs(x1 - x0) / (nx - 1), abs(y1 - y0) / (ny - 1), abs(z1 - z0) / (nz - 1)
    it = np.arange(nx * ny * nz)
    cx, cy, cz = it % nx, (it // nx) % ny, it // (nx * ny)
    pts = np.stack([x0 + cx * dx + 1e-12, y0 + cy * dy + 1e-12, z0 + cz * dz + 1e-12], axis=1)
    return pts, (nx, ny, nz)


def spherical_projection(pts, centre, field):
    """Result::getEHFields with projection on (Result.cpp:279-289, the
    rotation of Tools::fromProjection, Tools.cpp:291-301): the Cartesian
    field re-expressed in (r, theta, phi) components about the centre of the
    first sphere; returns (n, 3) complex."""
    d = pts - centre
    r = np.sqrt(np.sum(d * d, axis=1))
    safe = np.where(r > 0, r, 1.0)
    the = np.where(r > 0, np.arccos(np.clip(d[:, 2] / safe, -1, 1)), 0.0)
    phi = np.where(r > 0, np.arctan2(d[:, 1], d[:, 0]), 0.0)
    st, ct, sp, cp = np.sin(the), np.cos(the), np.sin(phi), np.cos(phi)
    fx, fy, fz = field[:, 0], field[:, 1], field[:, 2]
    return np.stack([st * cp * fx + st * sp * fy + ct * fz, ct * cp * fx + ct * sp * fy - st * fz,
                     cp * fy - sp * fx], axis=1)


def field_datasets(field, shape, prefix):
    """The datasets of one OutputGrid group: X, Y, Z real/imag and ABS
    (OutputGrid.cpp:160-210), indexed [i_x, i_y, i_z]."""
    nx, ny, nz = shape
    out = {}
    for c, ax in enumerate("XYZ"):
        v = field[:, c].reshape(nz, ny, nx).transpose(2, 1, 0)
        out[f"{prefix}/{ax}/real"] = np.ascontiguousarray(v.real)
        out[f"{prefix}/{ax}/imag"] = np.ascontiguousarray(v.imag)
    ab = np.sqrt(np.sum(np.abs(field) ** 2, axis=1)).reshape(nz, ny, nx).transpose(2, 1, 0)
    out[f"{prefix}/ABS/abs"] = np.ascontiguousarray(ab)
    return out


def fmt(v):
    """std::ostream's default formatting of a double (precision 6, %g)."""
    return f"{v:g}"


def write_response(prefix, lams, xs, sh):
    """The scan files of Simulation::scan_wavelengths (Simulation.cpp:620-671);
    xs rows hold (C_ext, C_sca, C_abs, C_sca_SH, C_abs_SH)."""
    files = {"_AbsorptionCS_FF.dat": xs[:, 0] - xs[:, 1], "_ScatteringCS_FF.dat": xs[:, 1]}
    if sh:
        files["_ScatteringCS_SH.dat"] = xs[:, 3]
        files["_AbsorptionCS_SH.dat"] = xs[:, 4]
    written = []
    for suffix, vals in files.items():
        path = prefix + suffix
        with open(path, "w") as f:
            for lam, v in zip(lams, vals):
                f.write(f"{fmt(lam)}\t{fmt(v)}\n")
        written.append(path)
    return written


def scan(case, solver_, batch=0):
    """Cross sections at every wavelength of the scan: one batched sweep when
    only the permittivities vary with the wavelength, a solve per wavelength
    when the susceptibilities vary too (the gold model)."""
    lams = case.wavelengths
    per = [case.at(l) for l in lams]
    batched = solver_.kind != "aca" and not any(s.material.varying_susceptibilities for s in case.spheres)
    if batched:
        solver_.set_cluster(per[0])
        eps = np.array([c.eps_r for c in per])
        eps_sh = np.array([c.eps_r_sh for c in per])
        out = solver_.sweep(lams, batch=batch, outputs=(), eps=eps, eps_sh=eps_sh)
        return out["xs"]
    xs = np.zeros((len(lams), 5))
    for i, c in enumerate(per):
        solver_.set_cluster(c)
        solver_.update()
        solver_.solve()
        cs = solver_.cross_sections()
        xs[i] = [cs["cext"], cs["csca"], cs["cabs"], cs["csca_sh"], cs["cabs_sh"]]
    return xs


def run_case(case, prefix, kind=None, context=None, batch=0):
    """Solves the case on the device and writes its outputs; returns the
    written paths and the computed data (xs for a scan, the fields for a
    field map).  kind None: "aca" when the case asks for ACA compression,
    else "lu32ir"."""
    from . import solver

    kind = kind or ("aca" if case.aca else "lu32ir")
    ctx = context or solver.Context()
    try:
        if case.output == "response":
            s = solver.Solver(case.at(case.wavelengths[0]), kind=kind, context=ctx)
            xs = scan(case, s, batch=batch)
            return write_response(prefix, case.wavelengths, xs, case.sh), xs
        c = case.cluster
        s = solver.Solver(c, kind=kind, context=ctx)
        s.update()
        s.solve()
        pts, shape = grid_points(case.grid)
        f = s.fields(pts)
        if case.projection:
            # Only the fields at omega are projected; the reference leaves the
            # second-harmonic fields at their zero default in that branch.
            f = dict(f, E=spherical_projection(pts, c.xyz[0], f["E"]), H=spherical_projection(pts, c.xyz[0], f["H"]))
            for k in ("E_sh", "H_sh"):
                if k in f:
                    f[k] = np.zeros_like(f[k])
        ff = dict(field_datasets(f["E"], shape, "Field_E"))
        ff.update(field_datasets(f["H"], shape, "Field_H"))
        np.savez(prefix + "_FF.npz", **ff)
        paths = [prefix + "_FF.npz"]
        if c.sh:
            shd = dict(field_datasets(f["E_sh"], shape, "Field_E"))
            shd.update(field_datasets(f["H_sh"], shape, "Field_H"))
            np.savez(prefix + "_SH.npz", **shd)
            paths.append(prefix + "_SH.npz")
        return paths, f
    finally:
        if context is None:
            ctx.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description="Solve an OPTIMET case file on the MI355X backend")
    ap.add_argument("case", help="the XML case file (with or without .xml)")
    ap.add_argument("--kind", default=None, choices=("lu32ir", "lu64", "gmres64", "aca"),
                    help="solver (default: aca when the case asks for ACA compression, else lu32ir)")
    ap.add_argument("--out", default=None, help="output prefix (default: the case path without .xml)")
    args = ap.parse_args(argv)
    path = args.case if args.case.endswith(".xml") else args.case + ".xml"
    case = read_case(path)
    prefix = args.out or os.path.splitext(path)[0]
    paths, _ = run_case(case, prefix, kind=args.kind)
    for p in paths:
        print(p)


if __name__ == "__main__":
    main()
