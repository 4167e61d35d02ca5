"""The CPU adjoint checker (oracle/adjoint.py) pinned before it is trusted:

* its steady residual vanishes at the CPU restatement's converged state
  (which the reference's own fixtures pin): a dropped or mis-signed term
  leaves a residual of the size of that term;
* its Jacobian times a random vector equals the central difference of the
  residual (an attribution slip in the coloured assembly fails this);
* the gradient it gives equals central differences of whole converged
  solves of the CPU restatement, for amuf, prsc and two inlet rows (with
  the east u raised by 1/nfy), for both objectives;
* dJ1/dprsc = 0 exactly (the scalar never enters u, v or p), and the
  adjoint and the forward derivative from the same Jacobian agree.
"""
import numpy as np
import pytest

import adjoint
import orc
from mgfvm.inputs import Problem, random_mask

NF = 32


def _case(tmp_path):
    j = (np.arange(NF) + 0.5) / NF
    prof = 1.0 + 0.3 * np.cos(np.pi * j) + 0.1 * np.sin(5.0 * np.pi * j)
    prob = Problem(nfx=NF, nfy=NF, ngrid=3, ig_obs=3, isolve=2, nvcycle=3000, nswpp=6,
                   convergence=1e-13, bc_xp=(float(prof.mean()), 0.0, 0.0, 0.0),
                   obstacles=random_mask(8, 0.85, 3))
    return prob, prof


def _solved(prob, prof, d):
    pin, ppat = prob.write(str(d))
    o = orc.Oracle(pin, ppat, inlet_u=prof)
    h, n, c = o.solve(prob.nvcycle, prob.convergence)
    assert c
    A = adjoint.from_oracle(o, prob)
    o.close()
    return A


@pytest.fixture(scope="module")
def base(tmp_path_factory):
    prob, prof = _case(None)
    A = _solved(prob, prof, tmp_path_factory.mktemp("adj"))
    K = A.jacobian()
    return prob, prof, A, K, {obj: A.solve_adjoint(obj, K) for obj in (0, 1)}


def test_residual_vanishes_at_converged_state(base):
    _, _, A, _, _ = base
    R = A.residual_at_state()
    for f, r in R.items():
        assert np.abs(r).max() < 1e-13, f


def test_jacobian_matches_central_differences(base):
    _, _, A, K, _ = base
    rng = np.random.default_rng(0)
    v = rng.standard_normal(A.n)

    def Rvec(x):
        args = list(A._base())
        for f, pos in (("u", 0), ("v", 1), ("p", 2), ("s", 3)):
            a = args[pos].copy()
            m = A.masks[f]
            a[m] = a[m] + x[A.col[f][m]]
            args[pos] = a
        return A._rowvec(A.residual(*args)).real

    eps = 1e-6
    fd = (Rvec(eps * v) - Rvec(-eps * v)) / (2 * eps)
    kv = K @ v
    assert np.abs(fd - kv).max() < 1e-8 * np.abs(kv).max()


def test_adjoint_equals_forward_derivative(base):
    _, _, A, K, res = base
    for obj in (0, 1):
        g, djdu = A.objective_grad(obj)
        for key, dth in (("amuf", {"amuf": 1.0}), ("prsc", {"prsc": 1.0})):
            fwd = g @ A.tangent(dth, K)
            adj = res[obj][3] if key == "amuf" else res[obj][4]
            assert abs(fwd - adj) <= 1e-10 * max(abs(fwd), 1e-300), (obj, key)
        e = np.zeros(A.nfy)
        e[5] = 1.0
        fwd = g @ A.tangent({"inlet": e, "uxp": 1.0 / A.nfy}, K) + djdu / A.nfy
        assert abs(fwd - res[obj][2][5]) <= 1e-10 * abs(fwd)


def test_pressure_drop_does_not_depend_on_prsc(base):
    assert base[4][0][4] == 0.0


def test_gradient_matches_central_differences_of_solves(base, tmp_path):
    prob, prof, A, _, res = base
    k = [0]

    def J(pr, pf):
        k[0] += 1
        return np.array(_solved(pr, pf, tmp_path / f"s{k[0]}").objectives())

    checks = []
    h = 1e-8
    fd = (J(prob.replace(amuf=prob.amuf + h), prof) - J(prob.replace(amuf=prob.amuf - h), prof)) / (2 * h)
    checks.append(("amuf", fd, [res[0][3], res[1][3]]))
    h = 1e-6
    fd = (J(prob.replace(prsc=prob.prsc + h), prof) - J(prob.replace(prsc=prob.prsc - h), prof)) / (2 * h)
    checks.append(("prsc", fd[1:], [res[1][4]]))
    for r in (0, 13):
        h = 1e-5
        e = np.zeros(NF)
        e[r] = 1.0
        ux = prob.bc_xp[0]
        fd = (J(prob.replace(bc_xp=(ux + h / NF, 0.0, 0.0, 0.0)), prof + h * e)
              - J(prob.replace(bc_xp=(ux - h / NF, 0.0, 0.0, 0.0)), prof - h * e)) / (2 * h)
        checks.append((f"inlet row {r + 1}", fd, [res[0][2][r], res[1][2][r]]))
    for name, fd, adj in checks:
        np.testing.assert_allclose(adj, fd, rtol=1e-6, err_msg=name)
