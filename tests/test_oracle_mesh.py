"""Pins the oracle's particle <-> mesh coupling to facts fixed by mathematics.

Deposition: volume fractions of hand-placed particles, conservation of the
deposited volume, the trilinear weights' zeroth and first moments (a
particle at a cell centre stays in that cell; the weighted cell centres give
back the particle position).  Interpolation: trilinear interpolation and the
centred gradient are exact on linear fields.  Chemistry
(xferMeshToParticleAndUpdateChem): the uptake half-steps converge (first
order) to the closed-form solution of the linear two-compartment ODE, amounts
are conserved between cell and fluid, the reaction settles on its
equilibrium k2 cA = kr2 cB cC with A+B and A+C conserved, volume growth is
linear in time at constant B amount (dilution), and the grown shapes keep
their geometric relations (4/3 pi r^3 = V for yeast; r/L fixed at the tip,
L fixed in the interior, pi r^2 L = V).  The shared pow is checked against
the C library.  Each check is chosen so that a wrong factor, sign, index or
axis order fails it.
"""
import math

import numpy as np
import pytest

NR, NI = 34, 35


def approx(v, rel):
    # no absolute floor: volumes and amounts here are ~1e-9 .. 1e-12
    return pytest.approx(v, rel=rel, abs=0)


def recs(oracle, pos, vol=1e-9, ctype=0):
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    r = np.zeros(len(pos), dtype=oracle.RECORD_DTYPE)
    r["pos"] = pos
    r["rdata"][:, 5] = vol
    r["rdata"][:, 0] = (3 * np.asarray(vol) / (4 * math.pi)) ** (1 / 3)
    r["idata"][:, 5] = ctype
    return r


def ulp_diff(a, b):
    ia = np.asarray(a, np.float64).view(np.int64)
    ib = np.asarray(b, np.float64).view(np.int64)
    return np.abs(ia - ib)


# ---- pow ------------------------------------------------------------------

def test_pow_agrees_with_libm(oracle):
    rng = np.random.default_rng(5)
    x = 10.0 ** rng.uniform(-30, 5, 20000) * rng.uniform(1, 2, 20000)
    for y in (1 / 3, 2 / 3, 0.5, 1.7, -2.3):
        mine = np.array([oracle.pow_(v, y) for v in x])
        d = ulp_diff(mine, [math.pow(v, y) for v in x])   # the C library's pow
        assert d.max() <= 1
        assert (d > 0).mean() < 5e-3


def test_pow_is_correctly_rounded_on_a_sample(oracle):
    """Against x^y to 50 digits: the double-double pow rounds correctly on
    every sampled argument (the C library's pow, < 1 ulp, need not)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    rng = np.random.default_rng(13)
    for _ in range(1500):
        x = float(10.0 ** rng.uniform(-30, 6) * rng.uniform(1, 2))
        y = float(rng.choice([1 / 3, 2 / 3, 0.5, 1.7, -2.3, rng.uniform(-4, 4)]))
        m = oracle.pow_(x, y)
        exact = (Decimal(x).ln() * Decimal(y)).exp()
        assert abs(Decimal(m) - exact) <= Decimal(math.ulp(m)) / 2, (x, y)


def test_pow_cube_root_path_gives_the_general_paths_bits(oracle):
    """pow(x, 1/3) and pow(x, 2/3) go through a cube root in double-double
    with a rounding test and fall back to the general double-double path
    when the test cannot decide: on 300,000 arguments over the whole normal
    range, on cell-sized volumes and on exact cubes, whatever it returns is
    that path's value, and it falls back rarely."""
    rng = np.random.default_rng(17)
    wide = 2.0 ** rng.uniform(-999, 999, 100000)
    vols = 10.0 ** rng.uniform(-14, -6, 100000) * rng.uniform(1, 2, 100000)
    t = np.round(rng.uniform(1, 2, 100000) * 2 ** 16) / 2 ** 16 * 2.0 ** rng.integers(-300, 300, 100000)
    cubes = t * t * t                                  # exact: 51-bit products
    fallbacks = 0
    for xs in (wide, vols, cubes):
        for k, y in ((1, 1 / 3), (2, 2 / 3)):
            for v in xs:
                f = oracle.pow_third(v, k)
                if f == 0.0:
                    fallbacks += 1
                    continue
                assert f == oracle.pow_dd(v, y), (v, k)
                assert oracle.pow_(v, y) == f
    assert fallbacks < 60                              # of 600,000 calls


def test_pow_exact_cases(oracle):
    rng = np.random.default_rng(6)
    x = 10.0 ** rng.uniform(-20, 5, 2000)
    # sqrt and squaring are correctly rounded; so is a near-exact pow
    assert all(oracle.pow_(v, 0.5) == math.sqrt(v) for v in x)
    assert all(oracle.pow_(v, 2.0) == v * v for v in x)
    assert all(oracle.pow_(v, 1.0) == v for v in x)
    assert oracle.pow_(8.0, 1 / 3) == 2.0


# ---- deposition ------------------------------------------------------------

def test_one_to_one_volume_fraction_known_cells(oracle):
    h = 0.25
    g = oracle.make_geom((0, 0, 0), (4 * h, 3 * h, 2 * h), (0, 0, 0), (4, 3, 2))
    cv = h ** 3
    pos = [[0.5 * h, 0.5 * h, 0.5 * h],          # cell (0, 0, 0)
           [3.2 * h, 1.0 * h, 1.7 * h],          # cell (3, 1, 1): on the j face
           [3.9 * h, 1.9 * h, 1.1 * h],          # cell (3, 1, 1)
           [1.5 * h, 2.5 * h, 0.2 * h]]          # cell (1, 2, 0), over-full
    vol = np.array([0.2, 0.1, 0.25, 1.5]) * cv
    r = recs(oracle, pos, vol)
    vf, out = oracle.volume_fraction(r, g, oracle.DEPOSIT_ONE_TO_ONE)
    assert out == 0
    want = np.ones((2, 3, 4))
    want[0, 0, 0] = 0.8
    want[1, 1, 3] = 1 - 0.35
    want[0, 2, 1] = 0.1
    np.testing.assert_allclose(vf, want, rtol=1e-14, atol=0)
    cnt, _ = oracle.count_per_cell(r, g)
    assert cnt[1, 1, 3] == 2 and cnt[0, 0, 0] == 1 and cnt.sum() == 4


def test_particles_outside_the_mesh_are_counted(oracle):
    g = oracle.make_geom((0, 0, 0), (1, 1, 1), (0, 0, 0), (4, 4, 4))
    r = recs(oracle, [[0.5, 0.5, 0.5], [0.5, 1.0, 0.5], [-1e-9, 0.2, 0.2]])
    _, out = oracle.volume_fraction(r, g, 0)
    assert out == 2


@pytest.mark.parametrize("scheme", [0, 1])
def test_deposit_conserves_volume_in_a_periodic_box(oracle, scheme):
    rng = np.random.default_rng(9)
    g = oracle.make_geom((0, 0, 0), (1, 2, 1.5), (1, 1, 1), (5, 8, 6))
    n = 3000
    vol = rng.uniform(1e-6, 1e-4, n)
    r = recs(oracle, rng.uniform(0, 1, (n, 3)) * [1, 2, 1.5], vol)
    m, out = oracle.deposit(r, g, scheme, 5, 1)
    assert out == 0
    np.testing.assert_allclose(m.sum(), vol.sum(), rtol=1e-13)


def test_trilinear_drops_only_beyond_walls(oracle):
    rng = np.random.default_rng(10)
    g = oracle.make_geom((0, 0, 0), (1, 1, 1), (0, 0, 0), (8, 8, 8))
    inner = recs(oracle, rng.uniform(1 / 16, 15 / 16, (500, 3)), 1e-5)
    m, _ = oracle.deposit(inner, g, 1, 5, 1)
    np.testing.assert_allclose(m.sum(), 500 * 1e-5, rtol=1e-13)
    edge = recs(oracle, [[0.01, 0.5, 0.5]], 1e-5)   # a quarter cell from x = 0
    m, _ = oracle.deposit(edge, g, 1, 5, 1)
    # l = 0.08 * 8 + 0.5 = 0.58: weight 0.42 goes to the ghost column i = -1
    np.testing.assert_allclose(m.sum(), 1e-5 * 0.58, rtol=1e-12)


def test_trilinear_weights_moments(oracle):
    rng = np.random.default_rng(11)
    n = (6, 5, 7)
    L = np.array([1.2, 1.0, 1.4])
    g = oracle.make_geom((0, 0, 0), L, (1, 1, 1), n)
    dx = L / n
    # a particle at a cell centre stays in that cell
    r = recs(oracle, [[(2.5) * dx[0], (3.5) * dx[1], (1.5) * dx[2]]], 1.0)
    m, _ = oracle.deposit(r, g, 1, 5, 1)
    assert m[0, 1, 3, 2] == approx(1.0, rel=1e-14)
    assert np.count_nonzero(m > 1e-14) == 1   # rounding of x * dxi aside
    # weighted cell centres give back the position (away from the wrap)
    for _ in range(20):
        x = (rng.uniform(1.5, [5.5, 4.5, 6.5])) * dx
        r = recs(oracle, [x], 1.0)
        m, _ = oracle.deposit(r, g, 1, 5, 1)
        k, j, i = np.meshgrid(*[np.arange(v) for v in n[::-1]], indexing="ij")
        cx = ((i + 0.5) * dx[0] * m[0]).sum()
        cy = ((j + 0.5) * dx[1] * m[0]).sum()
        cz = ((k + 0.5) * dx[2] * m[0]).sum()
        np.testing.assert_allclose([cx, cy, cz], x, rtol=1e-13)


# ---- interpolation ----------------------------------------------------------

def linear_field(n, L, a, b, lo=(0, 0, 0)):
    """a + b . x at every cell centre, one ghost layer, (nz+2, ny+2, nx+2)."""
    dx = np.asarray(L) / n
    c = [lo[d] + (np.arange(-1, n[d] + 1) + 0.5) * dx[d] for d in range(3)]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    return a + b[0] * x + b[1] * y + b[2] * z


def test_interpolation_exact_on_linear_fields(oracle):
    rng = np.random.default_rng(12)
    n = (5, 4, 6)
    L = (1.0, 0.8, 1.5)
    g = oracle.make_geom((0, 0, 0), L, (0, 0, 0), n)
    b = np.array([0.7, -1.3, 2.1])
    f = linear_field(n, L, 3.0, b)
    dx = np.asarray(L) / n
    for _ in range(50):
        x = rng.uniform(0, 1, 3) * L
        v = oracle.interp(x, f, g, 1)[0]
        assert v == approx(3.0 + b @ x, rel=1e-13)
        c = np.floor(x / dx)
        v1 = oracle.interp(x, f, g, 0)[0]
        assert v1 == approx(3.0 + b @ ((c + 0.5) * dx), rel=1e-13)


# ---- chemistry ---------------------------------------------------------------

def cell(vol, area, c, ctype=0, radius=None, L=None, position=0, n_bnds=1):
    rd = np.zeros(NR)
    rd[5] = vol
    rd[4] = area
    rd[0] = radius if radius is not None else (3 * vol / (4 * math.pi)) ** (1 / 3)
    rd[1] = L if L is not None else rd[0]
    rd[25:28] = c
    ip = np.zeros(NI, np.int32)
    ip[5] = ctype
    ip[19] = position
    ip[6] = n_bnds
    return rd, ip


def test_uptake_converges_to_the_two_compartment_solution(oracle):
    V, Vf, k1, kr1, t = 1e-9, 1e-8, 2.0, 1.0, 2e-4
    A = 4 * math.pi * (3 * V / (4 * math.pi)) ** (2 / 3)   # a yeast's area
    rd, ip = cell(V, A, (0.1, 0.0, 0.0))
    cp = [k1, 0, 0, kr1, 0, 0, 0, 0, 1, 1]
    lam = A * (k1 / Vf + kr1 / V)
    u0 = k1 * 1.0 - kr1 * 0.1
    exact = 0.1 + (A / V) * u0 * (1 - math.exp(-lam * t)) / lam
    res = {}
    for nl in (2000, 4000):
        out, err = oracle.chem_update(Vf, 1.0, rd, ip, (1.0, 0.0, 0.0), cp, t, nl)
        assert err == 0
        res[nl] = out[25]
        # the cell gains what the fluid loses (rdata[31] = amount taken)
        assert (out[25] - 0.1) * V == approx(-out[31], rel=1e-11)
    e1, e2 = res[2000] - exact, res[4000] - exact
    assert abs(e2 / e1 - 0.5) < 0.02          # first order
    assert 2 * res[4000] - res[2000] == approx(exact, rel=1e-6)


def test_reaction_reaches_equilibrium_with_conserved_sums(oracle):
    k2, kr2 = 40.0, 300.0
    rd, ip = cell(1e-9, 0.0, (1.0, 0.05, 0.02))
    cp = [0, k2, 0, 0, kr2, 0, 0, 0, 1, 1]
    out, err = oracle.chem_update(1e-8, 1.0, rd, ip, (0.0, 0.0, 0.0), cp, 2.0, 4000)
    assert err == 0
    a, b, c = out[25:28]
    assert a + b == approx(1.05, rel=1e-13)
    assert a + c == approx(1.02, rel=1e-13)
    assert k2 * a == approx(kr2 * b * c, rel=1e-10)
    np.testing.assert_array_equal(out[28:31], out[25:28])
    assert np.all(out[31:34] == 0)


def test_yeast_growth_is_linear_and_spherical(oracle):
    V0, cB, kv, dt = 1e-9, 0.3, 2.0, 0.5
    r0 = (3 * V0 / (4 * math.pi)) ** (1 / 3)
    rd, ip = cell(V0, 4 * math.pi * r0 ** 2, (0.0, cB, 0.0))
    cp = [0, 0, 0, 0, 0, 0, 0.0, kv, 1, 1]
    out, err = oracle.chem_update(1e-8, 1.0, rd, ip, (0.0, 0.0, 0.0), cp, dt, 7)
    assert err == 0
    V = out[5]
    assert V == approx(V0 * (1 + kv * cB * dt), rel=1e-14)
    assert out[26] * V == approx(cB * V0, rel=1e-14)     # B only diluted
    r = out[0]
    assert 4 / 3 * math.pi * r ** 3 == approx(V, rel=1e-14)
    assert out[4] == approx(4 * math.pi * r * r, rel=1e-15)
    assert out[22] == approx(kv * cB * V0, rel=1e-14)    # dV/dt
    assert out[21] == approx(2 * out[22] / r, rel=1e-14)


@pytest.mark.parametrize("position,r0,L0,rmax,lmax", [
    (0, 1e-4, 1e-3, 1.0, 1.0),      # tip below both limits: r and L scale
    (0, 1e-4, 1e-3, 1.0, 1e-3),     # tip at the length limit: r grows
    (0, 1e-4, 1e-3, 1e-4, 1.0),     # tip at the radius limit: L grows
    (3, 1e-4, 1e-3, 1.0, 1.0),      # interior segment: r grows, L fixed
])
def test_segment_growth_shapes(oracle, position, r0, L0, rmax, lmax):
    V0 = math.pi * r0 * r0 * L0
    rd, ip = cell(V0, 2 * math.pi * r0 * L0, (0.0, 0.4, 0.0), ctype=1, radius=r0, L=L0,
                  position=position)
    cp = [0, 0, 0, 0, 0, 0, 0.0, 3.0, rmax, lmax]
    out, err = oracle.chem_update(1e-8, 1.0, rd, ip, (0.0, 0.0, 0.0), cp, 0.1, 5)
    assert err == 0
    V, r, L = out[5], out[0], out[1]
    assert V == approx(V0 * (1 + 3.0 * 0.4 * 0.1), rel=1e-13)
    assert math.pi * r * r * L == approx(V, rel=1e-13)
    if position == 0 and rmax == 1.0 and lmax == 1.0:
        assert r / L == approx(r0 / L0, rel=1e-14)
    elif position == 3 or lmax < 1.0:
        assert L == L0 and r > r0
    else:
        assert r == r0 and L > L0
    if position == 3:
        assert out[4] == approx(2 * math.pi * r * L, rel=1e-15)


def test_segment_at_both_limits_does_not_grow(oracle):
    rd, ip = cell(1e-10, 1e-6, (0.0, 0.4, 0.0), ctype=1, radius=1e-4, L=1e-3, position=0)
    cp = [0, 0, 0, 0, 0, 0, 5.0, 3.0, 1e-4, 1e-3]
    out, err = oracle.chem_update(1e-8, 1.0, rd, ip, (0.0, 0.0, 0.0), cp, 0.1, 5)
    assert err == 0 and out[5] == 1e-10 and out[26] == 0.4


def test_uptake_empties_the_fluid_without_going_negative(oracle):
    V, A, Vf, f0 = 1e-9, 5e-6, 1e-8, 1e-3
    rd, ip = cell(V, A, (0.0, 0.0, 0.0))
    cp = [1e6, 0, 0, 0, 0, 0, 0, 0, 1, 1]
    out, err = oracle.chem_update(Vf, 1.0, rd, ip, (f0, 0.0, 0.0), cp, 1.0, 1)
    assert err == 0
    # the cell takes all the fluid holds and nothing more; the reference
    # undoes the overdraft by f -= d/Vf; f += d/Vf, which leaves f to within
    # one rounding of d/Vf
    taken = -out[31]
    assert out[25] * V == approx(taken, rel=1e-15)
    d1 = 0.5 * A * (1e6 * f0)
    assert abs(taken - f0 * Vf) <= 2 * math.ulp(d1 / Vf) * Vf


def test_negative_fluid_concentration_is_an_error(oracle):
    rd, ip = cell(1e-9, 5e-6, (0.0, 0.0, 0.0))
    _, err = oracle.chem_update(1e-8, 1.0, rd, ip, (-1e-3, 0, 0), [0] * 8 + [1, 1], 1.0, 1)
    assert err & 1
    _, err = oracle.chem_update(1e-8, 1.0, rd, ip, (-1e-16, 0, 0), [0] * 8 + [1, 1], 1.0, 1)
    assert err == 0


def test_mesh_to_particle_fluid_share_and_gradient(oracle):
    """Each particle draws on grid_vol * vf / npart of its cell's fluid; the
    gradient of species A at its cell is that of the linear field."""
    n, L = (4, 4, 4), (1.0, 1.0, 1.0)
    g = oracle.make_geom((0, 0, 0), L, (0, 0, 0), n)
    h = 0.25
    pos = [[0.1, 0.1, 0.1], [0.2, 0.15, 0.05], [0.6, 0.7, 0.8]]
    r = recs(oracle, pos, 1e-9)
    r["rdata"][:, 4] = 5e-6
    X = np.zeros((3, 6, 6, 6))
    b = np.array([1e-4, 2e-4, -3e-5])
    X[0] = linear_field(n, L, 1e-3, b)
    vf = np.full((6, 6, 6), 0.5)
    vf[1 + 3, 1 + 2, 1 + 2] = 0.8     # cell (i, j, k) = (2, 2, 3) of the third
    cp = [1e6, 0, 0, 0, 0, 0, 0, 0, 1, 1]
    err = oracle.mesh_to_particle(r, g, X, vf, cp, 1.0, 1, 0, 0)
    assert err == 0
    cellv = h ** 3
    fA = [1e-3 + b @ ((np.floor(np.array(p) / h) + 0.5) * h) for p in pos]
    share = [cellv * 0.5 / 2, cellv * 0.5 / 2, cellv * 0.8 / 1]
    for q in range(3):
        assert r["rdata"][q, 31] == approx(-fA[q] * share[q], rel=1e-13)
        np.testing.assert_allclose(r["rdata"][q, 18:21], b, rtol=1e-10)
