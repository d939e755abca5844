"""Case driver on the CPU: the XML reader with its quirks, the lattice
builders, the material models, the output grid and the scan file format.
The case files under tests/golden/cases are written in the reference's input
format (Reader.cpp); the expected values below come from what each format
rule fixes, not from running the reader twice."""
import math
import os

import numpy as np
import pytest

from optimet_amd import driver, materials

CASES = os.path.join(os.path.dirname(__file__), "golden", "cases")


def case(name):
    return driver.read_case(os.path.join(CASES, name + ".xml"))


def write(tmp_path, text, name="c.xml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


BASE = """<simulation><harmonics nmax="{nmax}" /></simulation>
<source type="planewave"><wavelength value="{lam}" /><propagation theta="{the}" phi="{phi}" />
<polarization Etheta.real="1" Etheta.imag="0" Ephi.real="0" Ephi.imag="0.5" />{sh}</source>
<geometry>{bg}{objects}</geometry>
<output type="response"><scan><wavelength initial="{li}" final="{lf}" stepsize="{ls}" /></scan></output>
"""
SPHERE = """<object type="sphere"><cartesian x="{x}" y="0" z="0" /><properties radius="{r}" />
<epsilon type="relative" value.real="4" value.imag="0.1" /><mu type="relative" value.real="1" value.imag="0" />
</object>"""


def text(nmax=3, lam=600, the=0, phi=0, sh="", bg="", objects=None, li=500, lf=700, ls=100):
    objects = objects if objects is not None else SPHERE.format(x=0, r=50)
    return BASE.format(nmax=nmax, lam=lam, the=the, phi=phi, sh=sh, bg=bg, objects=objects, li=li, lf=lf, ls=ls)


def test_units_and_source(tmp_path):
    """nm -> m (consFrnmTom), degrees -> radians, E_theta / E_phi, SH flag
    (Reader.cpp:590-605, 782-831)."""
    c = driver.read_case(write(tmp_path, text(lam=633, the=90, phi=45, sh='<SHsources condition="yes" />')))
    cl = c.cluster
    assert cl.wavelength == 633 * 1e-9
    assert cl.theta_k == pytest.approx(math.pi / 2, abs=1e-15) and cl.phi_k == pytest.approx(math.pi / 4, abs=1e-15)
    assert cl.etheta == 1 and cl.ephi == 0.5j and cl.sh
    assert cl.radius[0] == 50 * 1e-9 and cl.eps_r[0] == 4 + 0.1j and cl.mu_r[0] == 1
    assert cl.N == 2 * 3 * 5
    c2 = driver.read_case(write(tmp_path, text(sh='<SHsources condition="no" />')))
    assert not c2.cluster.sh


def test_spherical_position(tmp_path):
    obj = SPHERE.replace('<cartesian x="{x}" y="0" z="0" />', '<spherical rrr="200" the="1.2" phi="-0.4" />')
    c = driver.read_case(write(tmp_path, text(objects=obj.format(x=0, r=30))))
    r, t, p = 200e-9, 1.2, -0.4
    np.testing.assert_allclose(c.cluster.xyz[0], [r * math.sin(t) * math.cos(p), r * math.sin(t) * math.sin(p),
                                                  r * math.cos(t)], rtol=1e-15)


def test_background_quirk(tmp_path):
    """The background is read only when its type is NOT "relative"
    (Reader.cpp:84: strcmp(...) != 0)."""
    bg = ('<background type="{t}"><epsilon value.real="1.77" value.imag="0.01" />'
          '<mu value.real="1.1" value.imag="0" /></background>')
    c = driver.read_case(write(tmp_path, text(bg=bg.format(t="relative"))))
    assert c.cluster.bg_eps_r == 1 and c.cluster.bg_mu_r == 1
    c = driver.read_case(write(tmp_path, text(bg=bg.format(t="absolute"))))
    assert c.cluster.bg_eps_r == 1.77 + 0.01j and c.cluster.bg_mu_r == 1.1


@pytest.mark.parametrize("li,lf,ls,n", [(500, 1000, 20, 26), (600, 800, 150, 2), (1000, 1450, 3, 151),
                                        (500, 999.9, 100, 5)])
def test_scan_count(tmp_path, li, lf, ls, n):
    """steps = int(final - initial) / stepsize + 1 with C integer division, and
    lambda_i = initial + i (final - initial) / (steps - 1): the last wavelength
    is always `final`, whatever the step (Reader.cpp:886-891,
    Simulation.cpp:641-644)."""
    c = driver.read_case(write(tmp_path, text(li=li, lf=lf, ls=ls)))
    assert len(c.wavelengths) == n
    assert c.wavelengths[0] == li * 1e-9
    assert c.wavelengths[-1] == pytest.approx(lf * 1e-9, rel=1e-14)


def test_radius_scan_refused(tmp_path):
    """outputType 12 / 112 have no branch in Simulation::run (Simulation.cpp:45-63)."""
    t = text().replace('</scan>', '<radius initial="40" final="60" steps="3" /></scan>')
    with pytest.raises(driver.CaseError, match="radius scans"):
        driver.read_case(write(tmp_path, t))


def test_single_wavelength_scan_refused(tmp_path):
    with pytest.raises(driver.CaseError, match="single wavelength"):
        driver.read_case(write(tmp_path, text(li=600, lf=650, ls=100)))


def test_reader_errors(tmp_path):
    with pytest.raises(driver.CaseError, match="overlaps"):
        driver.read_case(write(tmp_path, text(objects=SPHERE.format(x=0, r=50) + SPHERE.format(x=90, r=40))))
    with pytest.raises(driver.CaseError, match="mu must be"):
        driver.read_case(write(tmp_path, text(objects=SPHERE.format(x=0, r=50).replace('type="relative" value.real="1"',
                                                                                       'type="x" value.real="1"'))))
    with pytest.raises(driver.CaseError, match="Unknown type for epsilon"):
        driver.read_case(write(tmp_path, text(objects=SPHERE.format(x=0, r=50).replace('epsilon type="relative"',
                                                                                       'epsilon type="Tabulated"'))))
    with pytest.raises(driver.CaseError, match="No scatterers"):
        driver.read_case(write(tmp_path, text(objects="")))
    with pytest.raises(driver.CaseError, match="only spheres"):
        driver.read_case(write(tmp_path, text(objects='<object type="arbitrary.shape" />')))
    with pytest.raises(driver.CaseError, match="At least 2 output steps"):
        t = text().replace('<output type="response"><scan><wavelength initial="500" final="700" stepsize="100" />'
                           '</scan></output>', '<output type="field"><grid><x min="0" max="1" steps="1" />'
                           '<y min="0" max="1" steps="2" /><z min="0" max="1" steps="2" /></grid></output>')
        driver.read_case(write(tmp_path, t))


def test_fixture_cases_parse():
    c = case("three_si_scan")
    assert c.output == "response" and c.sh and c.cluster.nobj == 3 and len(c.wavelengths) == 3
    c = case("gold_pair_scan")
    assert c.cluster.bg_eps_r == 1.77 and len(c.wavelengths) == 2 and c.wavelengths[-1] == pytest.approx(800e-9, rel=1e-15)
    c = case("two_spheres_field")
    assert c.output == "field" and not c.projection and c.cluster.bg_eps_r == 1  # type="relative": ignored
    assert c.cluster.ksippp[1] == 6.5e-18 and c.cluster.eps_r_sh[0] == 4.4 + 0.5j
    assert case("two_spheres_field_projected").projection


def test_cube_lattice_order():
    """The cube builder pushes the prototype, then for every lattice point but
    the first moves the last sphere there and pushes a fresh prototype: the
    spheres come out as lattice points 1..n-1 and then the prototype's own
    position (the origin) (Reader.cpp:123-178)."""
    c = case("cube_lattice_scan")
    d = 190 * 1e-9
    pts = [(d * i, d * j, d * k) for k in range(2) for j in range(2) for i in range(2)]
    np.testing.assert_array_equal(c.cluster.xyz, np.array(pts[1:] + pts[:1]))
    assert np.all(c.cluster.radius == 50 * 1e-9)


def test_zincblende_lattices(tmp_path):
    """GaAs lattices (Reader.cpp:183-540): a unit cell of 18 sites (14 of
    radius R, 4 of R / ratio), neighbouring cells sharing their faces, so the
    site count is that of a zinc-blende block: for n^3 cells (n+1)^3 corners
    + 3 n^2 (n+1) face centres + 4 n^3 interior sites; the origin site comes
    last, every site appears once."""
    obj = ('<object type="sphere"><properties radius="{r}" /><epsilon type="relative" value.real="4" '
           'value.imag="0" /><mu type="relative" value.real="1" value.imag="0" /></object>')
    for kind, n, ratio in (("GaAscube", 2, 1.7), ("GaAssurf", 2, 2.0)):
        struct = (f'<structure type="{kind}">' + obj.format(r=100) +
                  f'<properties No_cells="{n}" cell_size="1000" /></structure>')
        c = driver.read_case(write(tmp_path, text(objects=struct)))
        xyz, rad = c.cluster.xyz, c.cluster.radius
        nz = n if kind == "GaAscube" else 1
        corners = (n + 1) * (n + 1) * (nz + 1)
        faces = n * n * (nz + 1) + n * nz * (n + 1) + nz * n * (n + 1)
        inner = 4 * n * n * nz
        assert len(xyz) == corners + faces + inner
        assert len({tuple(np.round(p / 1e-12).astype(int)) for p in xyz}) == len(xyz)
        assert tuple(xyz[-1]) == (0.0, 0.0, 0.0) and rad[-1] == 100 * 1e-9
        assert np.sum(np.isclose(rad, 100e-9 / ratio, rtol=1e-15)) == inner
        assert np.sum(np.isclose(rad, 100e-9, rtol=1e-15)) == corners + faces


def test_silicon_model():
    """Tabulated n, k: exact at the table nodes, continuous across them,
    eps = (n + i k)^2, and the value at 2 omega is the value at lambda / 2."""
    lam_node = (0.25 + 55 * 0.01) * 1e-6
    e, esh, k3, k2, g = materials.silicon(lam_node)
    n, k = materials.SI_N[55], materials.SI_K[55]
    assert e == pytest.approx((n + 1j * k) ** 2, rel=1e-12)
    assert 13.0 < e.real < 14.0 and 0 < e.imag < 0.1  # silicon near 800 nm: n ~ 3.67
    assert (k3, k2, g) == (65e-19, 3.5e-19, 1.3e-19)
    for lam in (1.0123e-6, 1.2e-6, 1.4e-6):
        assert materials.silicon(lam)[1] == pytest.approx(materials.silicon(lam / 2)[0], rel=1e-12)
    node = 0.25 + 70 * 0.01
    left, right = materials.silicon(node * (1 - 1e-12) * 1e-6)[0], materials.silicon(node * (1 + 1e-12) * 1e-6)[0]
    assert left == pytest.approx(right, rel=1e-9)
    mid = materials.silicon((0.25 + 70.5 * 0.01) * 1e-6)[0]
    n5 = 0.5 * (materials.SI_N[70] + materials.SI_N[71])
    k5 = 0.5 * (materials.SI_K[70] + materials.SI_K[71])
    assert mid == pytest.approx((n5 + 1j * k5) ** 2, rel=1e-12)
    for bad in (0.49e-6, 1.46e-6):
        with pytest.raises(ValueError):
            materials.silicon(bad)


def test_gold_model():
    """The rational fit: eps at 2 omega equals eps at lambda / 2; gold near
    800 nm is a Drude metal (Re eps ~ -25, small positive Im eps); the
    hydrodynamic susceptibilities are -(a/4, b/2, d/8) (eps - 1) e / (m_e omega^2)."""
    e, esh, k3, k2, g = materials.gold(800e-9, 1.0, 1.0, 1.0)
    assert -30 < e.real < -20 and 0 < e.imag < 3
    assert esh == pytest.approx(materials.gold(400e-9, 1, 1, 1)[0], rel=1e-12)
    w = 2 * math.pi * materials.C0 / 800e-9
    unit = (e - 1) * materials.Q_E / (materials.M_E * w * w)
    assert k3 == pytest.approx(-unit / 4, rel=1e-12) and k2 == pytest.approx(-unit / 2, rel=1e-12)
    assert g == pytest.approx(-unit / 8, rel=1e-12)
    e2 = materials.gold(800e-9, 2.0 + 1j, 0, 0)
    assert e2[2] == pytest.approx((2.0 + 1j) * k3, rel=1e-12) and e2[0] == e


def test_scan_materials_per_wavelength():
    c = case("three_si_scan")
    for lam in c.wavelengths:
        cl = c.at(lam)
        assert cl.eps_r[0] == materials.silicon(lam)[0] and cl.eps_r_sh[2] == materials.silicon(lam)[1]
    g = case("gold_pair_scan")
    assert any(s.material.varying_susceptibilities for s in g.spheres)
    assert g.at(600e-9).ksippp[0] != g.at(800e-9).ksippp[0]


def test_grid_order():
    """OutputGrid: x fastest, then y, then z; +1e-12 m on every coordinate;
    step |max - min| / (n - 1) (OutputGrid.cpp:99-159)."""
    grid = [(-2e-7, 2e-7, 3), (0.0, 1e-7, 2), (5e-8, 1.5e-7, 2)]
    pts, shape = driver.grid_points(grid)
    assert shape == (3, 2, 2) and pts.shape == (12, 3)
    np.testing.assert_allclose(pts[0], [-2e-7 + 1e-12, 1e-12, 5e-8 + 1e-12], rtol=0, atol=1e-22)
    np.testing.assert_allclose(pts[1], [1e-12, 1e-12, 5e-8 + 1e-12], rtol=0, atol=1e-22)
    np.testing.assert_allclose(pts[3], [-2e-7 + 1e-12, 1e-7 + 1e-12, 5e-8 + 1e-12], rtol=0, atol=1e-22)
    np.testing.assert_allclose(pts[6], [-2e-7 + 1e-12, 1e-12, 1.5e-7 + 1e-12], rtol=0, atol=1e-22)
    field = np.zeros((12, 3), complex)
    field[:, 0] = np.arange(12) + 1j
    field[:, 2] = 2.0
    ds = driver.field_datasets(field, shape, "Field_E")
    assert ds["Field_E/X/real"].shape == (3, 2, 2)
    for it in range(12):
        i, j, k = it % 3, (it // 3) % 2, it // 6
        assert ds["Field_E/X/real"][i, j, k] == it and ds["Field_E/X/imag"][i, j, k] == 1
        assert ds["Field_E/ABS/abs"][i, j, k] == pytest.approx(math.sqrt(it * it + 1 + 4))


def test_spherical_projection():
    """A radial field projects onto (|E|, 0, 0), an azimuthal one onto (0, 0, |E|)."""
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(20, 3))
    c0 = np.array([0.1, -0.2, 0.3])
    d = pts - c0
    rhat = d / np.linalg.norm(d, axis=1)[:, None]
    a = rng.normal(size=20) + 1j * rng.normal(size=20)
    out = driver.spherical_projection(pts, c0, rhat * a[:, None])
    np.testing.assert_allclose(out[:, 0], a, rtol=1e-12)
    np.testing.assert_allclose(out[:, 1:], 0, atol=1e-12)
    phat = np.stack([-d[:, 1], d[:, 0], 0 * d[:, 0]], axis=1)
    phat /= np.linalg.norm(phat, axis=1)[:, None]
    out = driver.spherical_projection(pts, c0, phat * a[:, None])
    np.testing.assert_allclose(out[:, 2], a, rtol=1e-12)
    np.testing.assert_allclose(out[:, :2], 0, atol=1e-12)


def test_dat_format(tmp_path):
    """lambda<TAB>value per line with std::ostream's default precision 6
    (Simulation.cpp:659-666); C_abs is written as C_ext - C_sca."""
    lams = np.array([5e-7, 6.25e-7])
    xs = np.array([[3.0e-15, 1.0e-15, 0, 4.123456789e-25, 2.5e-26], [1.2345678e-14, 2.0e-15, 0, 1e-24, 7e-27]])
    paths = driver.write_response(str(tmp_path / "case"), lams, xs, sh=True)
    names = [os.path.basename(p) for p in paths]
    assert names == ["case_AbsorptionCS_FF.dat", "case_ScatteringCS_FF.dat", "case_ScatteringCS_SH.dat",
                     "case_AbsorptionCS_SH.dat"]
    assert open(paths[0]).read() == "5e-07\t2e-15\n6.25e-07\t1.03457e-14\n"
    assert open(paths[2]).read() == "5e-07\t4.12346e-25\n6.25e-07\t1e-24\n"
    assert open(paths[3]).read().splitlines()[1] == "6.25e-07\t7e-27"
    assert len(driver.write_response(str(tmp_path / "ff"), lams, xs, sh=False)) == 2
