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
