"""What the mathematics fixes about the CPU restatement, independently of the
reference's own outputs (closed forms, invariants, fixed points).  These also
pin the one part the reference cannot pin: obstacles below the obstacle
level (ig_obs < ngrid), where the reference reads uninitialised memory."""
import numpy as np
import pytest

import cases
import orc
from mgfvm.inputs import Problem, random_mask


def _oracle(tmp_path, prob, name="c"):
    pin, ppat = prob.write(str(tmp_path / name))
    return orc.Oracle(pin, ppat)


def test_hybrid_coefficients_closed_form(tmp_path):
    """coefu: aw = max(|F_w|, D) + F_w etc. with F the averaged half fluxes,
    ap = sum, walls doubled (V6:815-858)."""
    prob = cases.CASES["small_rb"]
    o = _oracle(tmp_path, prob)
    for k in range(3):
        o.vcycle(k)
    o.call("fluxes", 0)
    o.call("coefu", 0)
    cx, cy = o.get(0, "cx"), o.get(0, "cy")
    ny, nx = cx.shape
    dx, dy = prob.xl / prob.nfx, prob.yl / prob.nfy
    D = prob.amuf * (dy / dx)
    Dy = prob.amuf * (dx / dy)
    j, i = np.mgrid[1:ny - 1, 1:nx - 2]
    cxm = 0.5 * (cx[j, i] + cx[j, i - 1])
    cxp = 0.5 * (cx[j, i] + cx[j, i + 1])
    cym = 0.5 * (cy[j - 1, i] + cy[j - 1, i + 1])
    cyp = 0.5 * (cy[j, i] + cy[j, i + 1])
    aw = np.maximum(np.abs(cxm), D) + cxm
    ae = np.maximum(np.abs(cxp), D) - cxp
    as_ = np.maximum(np.abs(cym), Dy) + cym
    an = np.maximum(np.abs(cyp), Dy) - cyp
    an[-1, :] += Dy
    as_[0, :] += Dy
    ap = aw + ae + as_ + an
    np.testing.assert_allclose(o.get(0, "aw")[1:-1, 1:-2], aw, rtol=1e-14)
    np.testing.assert_allclose(o.get(0, "an")[1:-1, 1:-2], an, rtol=1e-14)
    np.testing.assert_allclose(o.get(0, "as")[1:-1, 1:-2], as_, rtol=1e-14)
    np.testing.assert_allclose(o.get(0, "ap")[1:-1, 1:-2], ap, rtol=1e-13)


def test_red_black_converges_to_linear_solve(tmp_path):
    """Many red-black sweeps on a diagonally dominant 5-point system reach
    the dense solution (the smoother's fixed point)."""
    prob = Problem(nfx=8, nfy=8, ngrid=1, ig_obs=1, isolve=2, nswpp=400, obstacles=[])
    o = _oracle(tmp_path, prob)
    rng = np.random.default_rng(0)
    shape = (10, 10)
    co = {n: np.zeros(shape) for n in ("aw", "ae", "as", "an")}
    for n in co:
        co[n][1:-1, 1:-1] = rng.uniform(0.1, 1.0, (8, 8))
    ap = sum(co.values()) + 0.5
    su = np.zeros(shape)
    su[1:-1, 1:-1] = rng.normal(size=(8, 8))
    for n, a in co.items():
        o.set(0, n, a)
    o.set(0, "ap", ap)
    o.set(0, "su", su)
    o.set(0, "pp", np.zeros(shape))
    o.call("rbgs_pp", 0)
    x = o.get(0, "pp")[1:-1, 1:-1]
    # dense system ap x - sum a_nb x_nb = su with zero boundary
    A = np.zeros((64, 64))
    b = su[1:-1, 1:-1].ravel()
    for jj in range(8):
        for ii in range(8):
            k = jj * 8 + ii
            A[k, k] = ap[jj + 1, ii + 1]
            for n, dj, di in (("aw", 0, -1), ("ae", 0, 1), ("as", -1, 0), ("an", 1, 0)):
                j2, i2 = jj + dj, ii + di
                if 0 <= j2 < 8 and 0 <= i2 < 8:
                    A[k, j2 * 8 + i2] = -co[n][jj + 1, ii + 1]
    np.testing.assert_allclose(x.ravel(), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_restriction_of_constants(tmp_path):
    """restv maps constant fields to the same constants (area-weighted faces,
    4-point cell means)."""
    prob = Problem(nfx=16, nfy=16, ngrid=2, ig_obs=2, obstacles=[])
    o = _oracle(tmp_path, prob)
    ny, nx = o.get(0, "u").shape
    for n, val in (("u", 0.75), ("v", -0.25), ("p", 3.0), ("sc", 0.5)):
        o.set(0, n, np.full((ny, nx), val))
    o.call("restv", 0)
    for n, val in (("u", 0.75), ("v", -0.25), ("p", 3.0), ("sc", 0.5)):
        np.testing.assert_allclose(o.get(1, n)[1:-1, 1:-1], val, rtol=1e-15)


def test_zero_flow_stays_zero(tmp_path):
    prob = Problem(nfx=32, nfy=32, ngrid=3, ig_obs=3, isolve=2, bc_xm=(0.0, 0.0, 0.0, 0.0),
                   bc_xp=(0.0, 0.0, 0.0, 0.0), guess=(0.0, 0.0, 0.0, 0.0),
                   obstacles=random_mask(8, 0.8, 5))
    o = _oracle(tmp_path, prob)
    hist, n, conv = o.solve(3, 1e-6)
    assert conv and n == 1 and hist[0] == 0.0
    for name in ("u", "v", "p", "sc"):
        assert not np.any(o.get(0, name))


def test_symmetric_channel(tmp_path):
    """Obstacle-free channel with symmetric data: u symmetric and v
    antisymmetric about the centreline (to round-off)."""
    prob = Problem(nfx=32, nfy=32, ngrid=3, ig_obs=3, isolve=1, nvcycle=20, obstacles=[])
    o = _oracle(tmp_path, prob)
    o.solve(20, 0.0)
    u = o.get(0, "u")[1:-1, :]
    v = o.get(0, "v")
    np.testing.assert_allclose(u, u[::-1, :], atol=1e-12)
    vin = v[1:-2, 1:-1]           # interior v faces 1..npy-3
    np.testing.assert_allclose(vin, -vin[::-1, :], atol=1e-12)


@pytest.mark.parametrize("name", ["cfg0_rb", "coarse_free"])
def test_converged_state_is_a_vcycle_fixed_point(tmp_path, name):
    """FAS consistency: once converged, another V-cycle changes nothing
    beyond round-off.  For coarse_free (obstacles only down to level 1, the
    rest defined by the library: coarsened obstacles, masked residuals and
    corrections) this is what pins the definition; without the residual mask
    the obstacle faces' penalty-balanced residual would drive corrections."""
    prob = cases.CASES[name].replace(nvcycle=3000, convergence=1e-11)
    o = _oracle(tmp_path, prob)
    hist, n, conv = o.solve(prob.nvcycle, prob.convergence)
    assert conv, hist[-5:]
    before = {k: o.get(0, k).copy() for k in ("u", "v", "p")}
    o.vcycle(n)
    for k, a in before.items():
        scale = max(np.abs(a).max(), 1.0)
        assert np.abs(o.get(0, k) - a).max() / scale < 1e-9, k


def test_coarse_obstacles_are_all_children_rule(tmp_path):
    """Below the obstacle level a cell is an obstacle iff all four children
    are (tag = 0 there)."""
    prob = cases.CASES["coarse_free"]
    o = _oracle(tmp_path, prob)
    for ig in range(prob.ig_obs, prob.ngrid):
        f = o.get(ig - 1, "tag") == 0
        c = o.get(ig, "tag") == 0
        want = f[1:-1:2, 1:-1:2] & f[2::2, 1:-1:2] & f[1:-1:2, 2::2] & f[2::2, 2::2]
        assert np.array_equal(c[1:-1, 1:-1], want), ig


def _random_system(o, seed):
    """A random diagonally dominant 5-point system on the 8x8 p' unknowns of
    level 0 (zero boundary); returns (coefficients, ap, su)."""
    rng = np.random.default_rng(seed)
    shape = (10, 10)
    co = {n: np.zeros(shape) for n in ("aw", "ae", "as", "an")}
    for n in co:
        co[n][1:-1, 1:-1] = rng.uniform(0.1, 1.0, (8, 8))
    ap = sum(co.values()) + 0.5
    su = np.zeros(shape)
    su[1:-1, 1:-1] = rng.normal(size=(8, 8))
    for n, a in co.items():
        o.set(0, n, a)
    o.set(0, "ap", ap)
    o.set(0, "su", su)
    return co, ap, su


def test_lexicographic_gs_converges_to_linear_solve(tmp_path):
    """The lexicographic Gauss-Seidel variant (not in the reference; pinned
    here by the mathematics): many sweeps reach the dense solution."""
    prob = Problem(nfx=8, nfy=8, ngrid=1, ig_obs=1, isolve=2, nswpp=200, obstacles=[])
    pin, ppat = prob.write(str(tmp_path / "lx"))
    o = orc.Oracle(pin, ppat, gs_lex=True)
    co, ap, su = _random_system(o, 1)
    o.set(0, "pp", np.zeros((10, 10)))
    o.call("lexgs_pp", 0)
    x = o.get(0, "pp")[1:-1, 1:-1]
    A = np.zeros((64, 64))
    b = su[1:-1, 1:-1].ravel()
    for jj in range(8):
        for ii in range(8):
            k = jj * 8 + ii
            A[k, k] = ap[jj + 1, ii + 1]
            for n, dj, di in (("aw", 0, -1), ("ae", 0, 1), ("as", -1, 0), ("an", 1, 0)):
                j2, i2 = jj + dj, ii + di
                if 0 <= j2 < 8 and 0 <= i2 < 8:
                    A[k, j2 * 8 + i2] = -co[n][jj + 1, ii + 1]
    np.testing.assert_allclose(x.ravel(), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    o.close()


def test_lexicographic_gs_sweep_order(tmp_path):
    """One lexicographic sweep from zero: the first point sees only its
    source (x = su / ap), each point's equation holds with its east and
    north neighbours at their old value (zero), and the last point's
    equation holds exactly afterwards (no neighbour of it changes later in
    the sweep).  A red-black sweep fails the second property."""
    prob = Problem(nfx=8, nfy=8, ngrid=1, ig_obs=1, isolve=2, nswpp=1, obstacles=[])
    pin, ppat = prob.write(str(tmp_path / "ls"))
    o = orc.Oracle(pin, ppat, gs_lex=True)
    co, ap, su = _random_system(o, 2)
    o.set(0, "pp", np.zeros((10, 10)))
    o.call("lexgs_pp", 0)
    x = o.get(0, "pp")
    assert x[1, 1] == su[1, 1] / ap[1, 1]
    # west / south neighbours at their new value, east / north at zero
    for j in range(1, 9):
        for i in range(1, 9):
            s = co["aw"][j, i] * x[j, i - 1] + co["as"][j, i] * x[j - 1, i]
            assert abs(ap[j, i] * x[j, i] - (s + su[j, i])) <= 1e-12 * (1 + abs(su[j, i]))
    r_last = ap[8, 8] * x[8, 8] - (co["aw"][8, 8] * x[8, 7] + co["as"][8, 8] * x[7, 8] + su[8, 8])
    assert abs(r_last) < 1e-12
    o.close()


def test_red_black_one_sweep_colour_order(tmp_path):
    """One red-black sweep from zero: the colour swept first ((i + j) even)
    sees only its source (x = su / ap), and the second colour's equations
    hold exactly afterwards (all their neighbours are of the first colour,
    which does not change again).  Swapping the colours fails both."""
    prob = Problem(nfx=8, nfy=8, ngrid=1, ig_obs=1, isolve=2, nswpp=1, obstacles=[])
    o = _oracle(tmp_path, prob)
    co, ap, su = _random_system(o, 3)
    o.set(0, "pp", np.zeros((10, 10)))
    o.call("rbgs_pp", 0)
    x = o.get(0, "pp")
    for j in range(1, 9):
        for i in range(1, 9):
            if (i + j) % 2 == 0:
                assert x[j, i] == su[j, i] / ap[j, i]
            else:
                s = (co["aw"][j, i] * x[j, i - 1] + co["ae"][j, i] * x[j, i + 1] +
                     co["as"][j, i] * x[j - 1, i] + co["an"][j, i] * x[j + 1, i])
                r = ap[j, i] * x[j, i] - (s + su[j, i])
                assert abs(r) <= 1e-12 * (1 + abs(su[j, i])), (j, i, r)


def test_masserror_stream_function_and_unit_face(tmp_path):
    """The continuity defect (V6:1096-1125) vanishes exactly for face
    velocities derived from a corner stream function (dx = dy = 1/16, integer
    stream function: every product exact), and a single perturbed u face
    moves su = -div by -/+ areax * delta in the two cells it separates."""
    prob = Problem(nfx=16, nfy=16, ngrid=1, ig_obs=1, obstacles=[])
    o = _oracle(tmp_path, prob)
    ny, nx = o.get(0, "u").shape
    ax, ay = o.get(0, "areax")[1, 1], o.get(0, "areay")[1, 1]
    assert ax == 1.0 / 16 and ay == 1.0 / 16
    rng = np.random.default_rng(4)
    psi = rng.integers(-8, 9, size=(ny, nx)).astype(float)
    u = np.zeros((ny, nx))
    v = np.zeros((ny, nx))
    u[1:, :] = (psi[1:, :] - psi[:-1, :]) / ax       # u[j,i] = (psi[j,i] - psi[j-1,i]) / ax
    v[:, 1:] = -(psi[:, 1:] - psi[:, :-1]) / ay      # v[j,i] = -(psi[j,i] - psi[j,i-1]) / ay
    o.set(0, "u", u)
    o.set(0, "v", v)
    o.set(0, "pp", np.full((ny, nx), 7.0))
    assert o.call("masserror", 0) == 0.0
    assert not np.any(o.get(0, "su")[1:-1, 1:-1])
    assert not np.any(o.get(0, "pp")[1:-1, 1:-1])
    delta = 0.375
    u[5, 6] += delta
    o.set(0, "u", u)
    err = o.call("masserror", 0)
    assert err == 2 * ax * delta
    su = o.get(0, "su")
    assert su[5, 6] == -ax * delta and su[5, 7] == ax * delta
    su[5, 6] = su[5, 7] = 0.0
    assert not np.any(su[1:-1, 1:-1])


def _cell_centres(n, h):
    return (np.arange(n) - 0.5) * h


@pytest.mark.parametrize("field", ["p", "sc"])
def test_cell_prolongation_reproduces_linear_corrections(tmp_path, field):
    """Bilinear 9/3/3/1 prolongation (V6:1894-2028): with the fine field zero
    (so the restricted fine value is zero), a coarse correction linear in x
    and y reaches every fine cell whose four coarse neighbours are interior
    as the same linear function at the fine cell centre.  A wrong neighbour
    or a 9/16 weight on the wrong cell fails it."""
    prob = Problem(nfx=16, nfy=16, ngrid=2, ig_obs=2, obstacles=[])
    o = _oracle(tmp_path, prob)
    nyf, nxf = o.get(0, field).shape
    nyc, nxc = o.get(1, field).shape
    hf, hc = 1.0 / 16, 1.0 / 8
    lin = lambda x, y: 1.0 + 0.75 * x - 2.5 * y
    yc, xc = np.meshgrid(_cell_centres(nyc, hc), _cell_centres(nxc, hc), indexing="ij")
    o.set(0, field, np.zeros((nyf, nxf)))
    o.set(1, field, lin(xc, yc))
    o.call("prol" + field, 0)
    got = o.get(0, field)
    yf, xf = np.meshgrid(_cell_centres(nyf, hf), _cell_centres(nxf, hf), indexing="ij")
    want = lin(xf, yf)
    # fine cells 2..nf-3 have all four coarse neighbours inside 1..nc-2
    np.testing.assert_allclose(got[2:-2, 2:-2], want[2:-2, 2:-2], rtol=0, atol=1e-14)


def test_face_prolongation_reproduces_linear_corrections(tmp_path):
    """Face prolongation of u (V6:1745-1820): 3/4-1/4 across rows at the
    coincident faces, halves onto the faces between them; a linear coarse
    correction reaches interior fine faces as the same linear function."""
    prob = Problem(nfx=16, nfy=16, ngrid=2, ig_obs=2, obstacles=[])
    o = _oracle(tmp_path, prob)
    nyf, nxf = o.get(0, "u").shape
    nyc, nxc = o.get(1, "u").shape
    hf, hc = 1.0 / 16, 1.0 / 8
    lin = lambda x, y: -0.5 + 1.25 * x + 3.0 * y
    yc, xc = np.meshgrid(_cell_centres(nyc, hc), np.arange(nxc) * hc, indexing="ij")
    o.set(0, "u", np.zeros((nyf, nxf)))
    o.set(1, "u", lin(xc, yc))
    o.call("prolu", 0)
    got = o.get(0, "u")
    yf, xf = np.meshgrid(_cell_centres(nyf, hf), np.arange(nxf) * hf, indexing="ij")
    want = lin(xf, yf)
    # fine rows 2..nyf-3, faces 3..nxf-5: both coarse rows and both coarse faces interior
    np.testing.assert_allclose(got[2:-2, 3:-4], want[2:-2, 3:-4], rtol=0, atol=1e-14)


def test_jacobi_uses_previous_sweep_only(tmp_path):
    """Two Jacobi sweeps from zero (V6:1231-1296): the first gives su / ap at
    every point, and the second point's value depends on the first sweep's
    neighbours only, so the result is the same whatever the visiting order
    (here: computed in reverse order with numpy); a Gauss-Seidel update in
    place fails it."""
    prob = Problem(nfx=8, nfy=8, ngrid=1, ig_obs=1, isolve=1, nswpp=1, obstacles=[])
    o = _oracle(tmp_path, prob)
    co, ap, su = _random_system(o, 7)
    o.set(0, "pp", np.zeros((10, 10)))
    o.call("jacobi_pp", 0)
    x1 = o.get(0, "pp").copy()
    np.testing.assert_array_equal(x1[1:-1, 1:-1], su[1:-1, 1:-1] / (ap[1:-1, 1:-1] + 1e-30))
    o.close()
    prob2 = prob.replace(nswpp=2)
    o = _oracle(tmp_path, prob2, "c2")
    co, ap, su = _random_system(o, 7)
    o.set(0, "pp", np.zeros((10, 10)))
    o.call("jacobi_pp", 0)
    x2 = o.get(0, "pp")
    want = np.zeros((10, 10))
    for j in range(8, 0, -1):
        for i in range(8, 0, -1):
            s = ((co["aw"][j, i] * x1[j, i - 1] + co["ae"][j, i] * x1[j, i + 1]) +
                 co["as"][j, i] * x1[j - 1, i]) + co["an"][j, i] * x1[j + 1, i]
            want[j, i] = (s + su[j, i]) / (ap[j, i] + 1e-30)
    np.testing.assert_allclose(x2[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-14, atol=0)
    o.close()
