"""Forward sensitivities (libmgfvm_tangent.so): the value fields equal the
FP64 library's bit for bit, and the derivatives of the objectives along
amuf, prsc and single inlet rows agree with central finite differences of
FP64 solves (tolerance 1e-5 relative, the gradient check the benchmark's
differentiation workload states)."""
import os
import sys

import numpy as np
import pytest

from mgfvm.inputs import Problem, Rng, random_mask

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(u_xp):
    """64x64 channel, 16x16 obstacle base (Bernoulli porosity 0.85), Re 100,
    red-black 2/6, plug east boundary at the mean inlet velocity (so inflow
    and outflow balance), converged to sum|div| < 1e-12."""
    return Problem(nfx=64, nfy=64, ngrid=3, ig_obs=3, nvcycle=3000, convergence=1e-12,
                   isolve=2, nswpm=2, nswpp=6, amuf=0.01, bc_xp=(u_xp, 0.0, 0.0, 0.0),
                   obstacles=random_mask(16, 0.85, 11))


def _inlet(nfy, seed=3):
    y = (np.arange(nfy) + 0.5) / nfy
    rng = Rng(seed)
    return 6.0 * y * (1.0 - y) + np.array([0.05 * rng.normal() for _ in range(nfy)])


def _objectives(s, yl):
    u, p, sc = s.get(0, "u"), s.get(0, "p"), s.get(0, "sc")
    npy, npx = u.shape
    ax = yl / (npy - 2)
    j1 = p[1:-1, 1].sum() / (npy - 2) - p[1:-1, npx - 2].sum() / (npy - 2)
    j2 = 0.0
    for j in range(1, npy - 1):
        j2 += (ax * u[j, npx - 2]) * sc[j, npx - 2]
    return np.array([j1, j2])


def _fp64_J(tmp_path, prob, inlet, tag):
    import mgfvm
    pin, ppat = prob.write(str(tmp_path / tag))
    s = mgfvm.Solver(pin, ppat, inlet_u=inlet)
    _, n, conv = s.solve()
    assert conv, tag
    J = _objectives(s, prob.yl)
    s.close()
    return J


def test_tangent_values_and_finite_differences(tmp_path):
    import mgfvm
    from mgfvm.tangent import TangentBatch
    inlet = _inlet(64)
    prob = _problem(float(inlet.mean()))
    pin, ppat = prob.write(str(tmp_path / "base"))
    rows = [5, 20, 40]
    seeds = [{"d_amuf": 1.0}, {"d_prsc": 1.0}]
    for r in rows:
        e = np.zeros(prob.nfy)
        e[r - 1] = 1.0
        seeds.append({"d_inlet_u": e, "d_u_xp": 1.0 / prob.nfy})
    _say(f"[gpu] tangent batch of {len(seeds)}")
    tb = TangentBatch([(pin, ppat)] * len(seeds), seeds, inlet_u=inlet)
    ncyc, conv = tb.solve()
    assert all(conv) and len(set(ncyc)) == 1

    s = mgfvm.Solver(pin, ppat, inlet_u=inlet)
    _, n, c = s.solve()
    assert c and n == ncyc[0]
    for name in ("u", "v", "p", "sc"):
        v, _ = tb.get(0, 0, name)
        assert np.array_equal(v.view(np.uint64), s.get(0, name).view(np.uint64)), name
    J0 = _objectives(s, prob.yl)
    s.close()
    Jt, _ = tb.objectives(0)
    np.testing.assert_allclose(Jt, J0, rtol=1e-13)

    # central steps: small enough relative to each input that no face of the
    # hybrid scheme changes branch (|F| vs amuf * area / spacing), large
    # enough that the round-off of J (~1e-13 relative) stays below 1e-6
    checks = []
    for k, sd in enumerate(seeds):
        _, dJ = tb.objectives(k)
        h = 1e-8 if "d_amuf" in sd else (1e-6 if "d_prsc" in sd else 1e-5)
        if "d_amuf" in sd:
            jp = _fp64_J(tmp_path, prob.replace(amuf=prob.amuf + h), inlet, f"amuf_p")
            jm = _fp64_J(tmp_path, prob.replace(amuf=prob.amuf - h), inlet, f"amuf_m")
        elif "d_prsc" in sd:
            jp = _fp64_J(tmp_path, prob.replace(prsc=prob.prsc + h), inlet, f"prsc_p")
            jm = _fp64_J(tmp_path, prob.replace(prsc=prob.prsc - h), inlet, f"prsc_m")
        else:
            e = sd["d_inlet_u"]
            up = (prob.bc_xp[0] + h / prob.nfy, 0.0, 0.0, 0.0)
            um = (prob.bc_xp[0] - h / prob.nfy, 0.0, 0.0, 0.0)
            jp = _fp64_J(tmp_path, prob.replace(bc_xp=up), inlet + h * e, f"in{k}_p")
            jm = _fp64_J(tmp_path, prob.replace(bc_xp=um), inlet - h * e, f"in{k}_m")
        fd = (jp - jm) / (2 * h)
        checks.append((k, dJ, fd))
        _say(f"[gpu] direction {k}: tangent {dJ} fd {fd}")
    tb.close()
    for k, dJ, fd in checks:
        for q in range(2):
            if abs(fd[q]) < 1e-9 and abs(dJ[q]) < 1e-9:
                continue          # both zero (J1 does not depend on prsc)
            assert abs(dJ[q] - fd[q]) <= REL_TOL * max(abs(fd[q]), 1e-12), (k, q, dJ[q], fd[q])


def _fd_case(tmp_path, prob, inlet, kind, h, e=None):
    if kind == "amuf":
        a, b = prob.replace(amuf=prob.amuf + h), prob.replace(amuf=prob.amuf - h)
        ia = ib = inlet
    elif kind == "prsc":
        a, b = prob.replace(prsc=prob.prsc + h), prob.replace(prsc=prob.prsc - h)
        ia = ib = inlet
    else:
        a = prob.replace(bc_xp=(prob.bc_xp[0] + h / prob.nfy, 0.0, 0.0, 0.0))
        b = prob.replace(bc_xp=(prob.bc_xp[0] - h / prob.nfy, 0.0, 0.0, 0.0))
        ia, ib = inlet + h * e, inlet - h * e
    return (_fp64_J(tmp_path, a, ia, f"{kind}_p") - _fp64_J(tmp_path, b, ib, f"{kind}_m")) / (2 * h)


def test_gradients_256_against_finite_differences(tmp_path):
    """The differentiation workload at full size (256x256, 5 levels,
    porosity-0.85 mask on a 32x32 base, Re 100, noisy parabolic inlet):
    mgfvm.tangent.gradients for inlet rows 1 and 128, amuf and prsc against
    central differences of FP64 solves."""
    from mgfvm.inputs import sensitivity_256
    from mgfvm.tangent import gradients
    prob, inlet = sensitivity_256(1)
    inlet = np.array(inlet)
    pin, ppat = prob.write(str(tmp_path / "base"))
    _say("[gpu] gradients 256x256")
    J, din, dmu, dpr, conv = gradients(pin, ppat, inlet_u=inlet, rows=[1, 128])
    assert conv
    cases = [("amuf", 1e-8, None, dmu), ("prsc", 1e-6, None, dpr)]
    for k, r in enumerate((1, 128)):
        e = np.zeros(prob.nfy)
        e[r - 1] = 1.0
        cases.append(("inlet", 1e-5, e, din[:, k]))
    for kind, h, e, d in cases:
        fd = _fd_case(tmp_path, prob, inlet, kind, h, e)
        _say(f"[gpu] 256^2 {kind}: tangent {d} fd {fd}")
        for q in range(2):
            if abs(fd[q]) < 1e-9 and abs(d[q]) < 1e-9:
                continue
            assert abs(d[q] - fd[q]) <= REL_TOL * abs(fd[q]), (kind, q, d[q], fd[q])


def test_torch_function_backward(tmp_path):
    """torch.autograd over the steady objectives: backward gives g . dJ/dtheta
    with the forward-mode Jacobian of the same solve."""
    import torch
    from mgfvm.tangent import steady_objectives, torch_function
    inlet = _inlet(64)
    prob = _problem(float(inlet.mean()))
    F = torch_function()
    x = torch.tensor(inlet, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(prob.amuf, dtype=torch.float64, requires_grad=True)
    pr = torch.tensor(prob.prsc, dtype=torch.float64, requires_grad=True)
    _say("[gpu] torch autograd over the steady objectives")
    J = F.apply(x, mu, pr, prob)
    (J[0] + 3.0 * J[1]).backward()
    J2, din, dmu, dpr, conv = steady_objectives(inlet, prob.amuf, prob.prsc, prob)
    np.testing.assert_allclose(J.detach().numpy(), J2, rtol=0, atol=0)
    w = np.array([1.0, 3.0])
    np.testing.assert_allclose(x.grad.numpy(), w @ din, rtol=1e-14)
    np.testing.assert_allclose(mu.grad.item(), w @ dmu, rtol=1e-14)
    np.testing.assert_allclose(pr.grad.item(), w @ dpr, rtol=1e-14)
    # float hyperparameters get no gradient slot; a device inlet keeps its device
    xd = torch.tensor(inlet, dtype=torch.float64, device="cuda:0", requires_grad=True)
    Jd = F.apply(xd, prob.amuf, prob.prsc, prob)
    assert Jd.device == xd.device
    Jd[0].backward()
    assert xd.grad.device == xd.device
    np.testing.assert_allclose(xd.grad.cpu().numpy(), din[0], rtol=1e-14)


def test_c_gradient_entry_matches_batched_directions(tmp_path):
    """mgf_tangent_gradient (the library's own gradient call: every inlet
    row, amuf and prsc, in batches of max_batch directions) gives the bits
    of the same directions solved as one Python-built batch, whatever the
    batch size; rows are mass-consistent directions (east u += 1/nfy)."""
    from mgfvm.tangent import gradient_c, gradients
    nf = 32
    inlet = _inlet(nf, seed=5)
    prob = Problem(nfx=nf, nfy=nf, ngrid=3, ig_obs=3, nvcycle=3000, convergence=1e-12,
                   isolve=2, nswpm=2, nswpp=6, amuf=0.01,
                   bc_xp=(float(inlet.mean()), 0.0, 0.0, 0.0),
                   obstacles=random_mask(8, 0.85, 4))
    pin, ppat = prob.write(str(tmp_path / "g"))
    _say("[gpu] mgf_tangent_gradient vs batched directions")
    ref = gradients(pin, ppat, inlet_u=inlet)
    assert ref[4]
    for mb in (0, 7):
        got = gradient_c(pin, ppat, inlet_u=inlet, max_batch=mb)
        assert got[4], mb
        for a, b in zip(got[:4], ref[:4]):
            assert np.array_equal(np.asarray(a), np.asarray(b)), mb
