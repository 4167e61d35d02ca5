"""The reverse-mode adjoint (include/mgfvm_adjoint.h) through the C ABI,
against the CPU adjoint checker (oracle/adjoint.py: complex-step Jacobian,
sparse direct solve; pinned by tests/test_oracle_adjoint.py), against the
forward-mode sensitivity library on every direction of the 256^2
differentiation workload, against central differences of whole GPU
solves, and on a reduced batch of the 512^2 batched-adjoint workload.

Tolerances: the transposed Jacobian is the same arithmetic up to rounding
(1e-13 relative to its largest entry); gradients are GMRES solutions to a
relative residual of 1e-11, so they agree with the direct solve and with
the forward-mode derivative (itself an iteration converged to the primal's
tolerance) to 1e-8 relative (to the largest component of each gradient);
central differences (step 1e-8 amuf, 1e-6 prsc, 1e-5 inlet, no hybrid
branch changes) to 1e-5.
"""
import sys
import tempfile

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-8


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case32():
    from mgfvm.inputs import Problem, random_mask
    nf = 32
    j = (np.arange(nf) + 0.5) / nf
    prof = 1.0 + 0.3 * np.cos(np.pi * j) + 0.1 * np.sin(5.0 * np.pi * j)
    prob = Problem(nfx=nf, nfy=nf, ngrid=3, ig_obs=3, isolve=2, nvcycle=3000, nswpp=6,
                   convergence=1e-13, bc_xp=(float(prof.mean()), 0.0, 0.0, 0.0),
                   obstacles=random_mask(8, 0.85, 3))
    return prob, prof


def _case64_coarse_obstacles():
    """64^2, 4 levels, obstacles on a 32^2 base (coarsened below it), dx != dy"""
    from mgfvm.inputs import Problem, random_mask
    nf = 64
    j = (np.arange(nf // 2) + 0.5) / (nf // 2)
    prof = 6.0 * j * (1.0 - j)
    prob = Problem(nfx=nf, nfy=nf // 2, ngrid=3, ig_obs=2, isolve=2, nvcycle=3000, nswpp=6,
                   xl=2.0, yl=0.5, amuf=0.005, convergence=1e-12,
                   bc_xp=(float(prof.mean()), 0.0, 0.0, 0.0),
                   obstacles=[r for r in random_mask(32, 0.85, 12) if r[3] <= 16])
    return prob, prof


def _pair(prob, prof, d, precision=0):
    import mgfvm
    pin, ppat = prob.write(str(d))
    s = mgfvm.Solver(pin, ppat, inlet_u=prof, precision=precision)
    o = orc.Oracle(pin, ppat, inlet_u=prof)
    return s, o, pin, ppat


def _close(a, b, rtol, what):
    a, b = np.atleast_1d(np.asarray(a, float)), np.atleast_1d(np.asarray(b, float))
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: relative error {err:.3e}"


def test_transposed_jacobian_matches_oracle(tmp_path):
    import adjoint as oadj
    from mgfvm import adjoint as gadj
    for name, (prob, prof) in (("32", _case32()), ("64x32", _case64_coarse_obstacles())):
        _say(f"[gpu] K^T vs oracle: {name}")
        s, o, _, _ = _pair(prob, prof, tmp_path / name)
        h, n, c = s.solve()
        ho, no, co = o.solve(prob.nvcycle, prob.convergence)
        assert c and co and n == no
        A = oadj.from_oracle(o, prob)
        K = A.jacobian()
        npx, npy = o.dims(0)
        rng = np.random.default_rng(1)
        lam = rng.standard_normal(A.n)
        grids = []
        for f in ("u", "v", "p", "s"):
            g = np.zeros((npy, npx))
            m = A.masks[f]
            g[m] = lam[A.col[f][m]]
            grids.append(g.ravel())
        out = gadj.apply(s, 0, 0, np.concatenate(grids))
        n0 = npx * npy
        got = np.zeros(A.n)
        for k, f in enumerate(("u", "v", "p", "s")):
            m = A.masks[f]
            got[A.col[f][m]] = out[k * n0:(k + 1) * n0].reshape(npy, npx)[m]
            # nothing outside the unknowns
            assert np.all(out[k * n0:(k + 1) * n0].reshape(npy, npx)[~m] == 0.0)
        _close(got, K.T @ lam, 1e-13, "K^T lambda")
        # the preconditioner is a linear map
        v1 = rng.standard_normal(4 * n0) * np.concatenate([g != 0 for g in grids])
        v2 = rng.standard_normal(4 * n0) * np.concatenate([g != 0 for g in grids])
        m1, m2 = gadj.apply(s, 0, 1, v1), gadj.apply(s, 0, 1, v2)
        m12 = gadj.apply(s, 0, 1, 2.0 * v1 - 3.0 * v2)
        _close(m12, 2.0 * m1 - 3.0 * m2, 1e-12, "V-cycle linearity")
        s.close()
        o.close()


@pytest.mark.parametrize("which", ["32", "64x32"])
def test_gradients_match_oracle_adjoint(tmp_path, which):
    import adjoint as oadj
    from mgfvm import adjoint as gadj
    prob, prof = _case32() if which == "32" else _case64_coarse_obstacles()
    _say(f"[gpu] adjoint gradients vs oracle: {which}")
    s, o, _, _ = _pair(prob, prof, tmp_path / which)
    h, n, c = s.solve()
    o.solve(prob.nvcycle, prob.convergence)
    A = oadj.from_oracle(o, prob)
    K = A.jacobian()
    for obj in (0, 1):
        _, J, din, dmu, dpr = A.solve_adjoint(obj, K)
        gi, ga, gp = gadj.adjoint(s, obj)
        assert abs(gadj.objectives(s)[obj] - J) <= 1e-12 * abs(J)
        _close(gi, din, GRAD_RTOL, f"obj {obj} inlet")
        _close(ga, dmu, GRAD_RTOL, f"obj {obj} amuf")
        if obj == 1:
            _close(gp, dpr, GRAD_RTOL, "obj 1 prsc")
        else:
            assert gp == 0.0           # sc never enters u, v or p
        lu, lv, lc, ls = gadj.state(s)
        if obj == 0:
            assert np.all(ls == 0.0)   # J1 does not depend on sc: lambda_s = 0
    s.close()
    o.close()


def test_differentiation_workload_all_directions_match_forward_mode(tmp_path):
    """The 256^2 differentiation workload: the two adjoint solves give every
    one of the 258 directions (256 inlet rows, amuf, prsc) of both
    objectives, equal to the forward-mode library's 258-direction batch."""
    import mgfvm
    from mgfvm import adjoint as gadj
    from mgfvm.inputs import sensitivity_256
    from mgfvm.tangent import gradient_c
    p, inlet = sensitivity_256(1)
    pin, ppat = p.write(str(tmp_path / "g256"))
    _say("[gpu] 256^2: forward-mode 258 directions")
    J, gin, gmu, gpr, conv = gradient_c(pin, ppat, inlet_u=np.array(inlet))
    assert conv
    _say("[gpu] 256^2: reverse mode, both objectives")
    with mgfvm.Solver(pin, ppat, inlet_u=np.array(inlet)) as s:
        h, n, c = s.solve()
        assert c
        for obj in (0, 1):
            r = gadj.batch_adjoint(s, obj)[0]
            Ja, gi, ga, gp, it, rr = r
            assert rr <= 1e-10 and it > 0
            assert abs(Ja - J[obj]) <= 1e-12 * abs(J[obj])
            _close(gi, gin[obj], GRAD_RTOL, f"obj {obj} inlet rows")
            _close(ga, gmu[obj], GRAD_RTOL, f"obj {obj} amuf")
            if obj == 1:
                _close(gp, gpr[obj], GRAD_RTOL, "obj 1 prsc")


def test_gradient_matches_central_differences(tmp_path):
    """amuf, prsc and two inlet rows of the 256^2 workload against central
    differences of FP64 GPU solves converged to sum |div| < 1e-12."""
    import mgfvm
    from mgfvm import adjoint as gadj
    from mgfvm.inputs import sensitivity_256
    p, inlet = sensitivity_256(1)
    inlet = np.array(inlet)
    nfy = p.nfy
    k = [0]

    def J(pr, prof):
        k[0] += 1
        pin, ppat = pr.write(str(tmp_path / f"fd{k[0]}"))
        with mgfvm.Solver(pin, ppat, inlet_u=prof) as s:
            h, n, c = s.solve()
            assert c
            return gadj.objectives(s)

    pin, ppat = p.write(str(tmp_path / "base"))
    _say("[gpu] 256^2 adjoint vs central differences")
    with mgfvm.Solver(pin, ppat, inlet_u=inlet) as s:
        s.solve()
        adj = {obj: gadj.adjoint(s, obj) for obj in (0, 1)}
    h = 1e-8
    fd = (J(p.replace(amuf=p.amuf + h), inlet) - J(p.replace(amuf=p.amuf - h), inlet)) / (2 * h)
    for obj in (0, 1):
        _close(adj[obj][1], fd[obj], 1e-5, f"amuf obj {obj}")
    h = 1e-6
    fd = (J(p.replace(prsc=p.prsc + h), inlet) - J(p.replace(prsc=p.prsc - h), inlet)) / (2 * h)
    _close(adj[1][2], fd[1], 1e-5, "prsc obj 1")
    for r in (0, 127):
        h = 1e-5
        e = np.zeros(nfy)
        e[r] = 1.0
        ux = p.bc_xp[0]
        fd = (J(p.replace(bc_xp=(ux + h / nfy, 0.0, 0.0, 0.0)), inlet + h * e)
              - J(p.replace(bc_xp=(ux - h / nfy, 0.0, 0.0, 0.0)), inlet - h * e)) / (2 * h)
        for obj in (0, 1):
            _close(adj[obj][0][r], fd[obj], 1e-5, f"inlet row {r + 1} obj {obj}")


def test_batched_adjoint_workload_reduced(tmp_path):
    """Four problems of the 512^2 batched-adjoint workload (6 levels, random
    masks with porosity in [0.7, 0.95], Re in [20, 300]) in one batch: the
    batched adjoint of each equals the single-problem adjoint, and agrees
    with forward-mode directions (amuf and two inlet rows) of the
    sensitivity library, both on FP64 primal solves; on a mixed-precision
    primal solve the adjoint gradient is within 1e-5 of the FP64 one."""
    import mgfvm
    from mgfvm import adjoint as gadj
    from mgfvm.inputs import adjoint_batch
    from mgfvm.tangent import TangentBatch
    probs = [q.replace(convergence=1e-10) for q in adjoint_batch(8)]
    pairs = [q.write(str(tmp_path / f"p{k}")) for k, q in enumerate(probs)]
    _say("[gpu] 512^2 batch: FP64 primal + batched adjoint")
    with mgfvm.Batch(pairs) as bt:
        nc, cv = bt.solve()
        res = gadj.batch_adjoint(bt, 0)
    idx = [k for k in range(len(probs)) if cv[k]][:4]
    assert len(idx) == 4
    for k in range(len(probs)):
        if cv[k]:
            assert res[k][4] > 0 and res[k][5] <= 1e-10
        else:
            assert res[k][4] == -1 and np.isnan(res[k][2])
    _say("[gpu] 512^2: single-problem adjoints and forward-mode directions")
    for k in idx[:2]:
        with mgfvm.Solver(*pairs[k]) as s:
            s.solve()
            gi, ga, gp = gadj.adjoint(s, 0)
        _close(gi, res[k][1], 1e-12, f"batch vs single inlet {k}")
        _close(ga, res[k][2], 1e-12, f"batch vs single amuf {k}")
        nfy = probs[k].nfy
        e1, e2 = np.zeros(nfy), np.zeros(nfy)
        e1[100] = 1.0
        e2[300] = 1.0
        seeds = [{"d_amuf": 1.0}, {"d_inlet_u": e1, "d_u_xp": 1.0 / nfy},
                 {"d_inlet_u": e2, "d_u_xp": 1.0 / nfy}]
        tb = TangentBatch([pairs[k]] * 3, seeds)
        tb.solve()
        d = [tb.objectives(q)[1][0] for q in range(3)]
        tb.close()
        scale = max(abs(res[k][2]), 1e-300)
        assert abs(d[0] - res[k][2]) <= GRAD_RTOL * scale, (k, d[0], res[k][2])
        gscale = np.abs(res[k][1]).max()
        assert abs(d[1] - res[k][1][100]) <= GRAD_RTOL * gscale
        assert abs(d[2] - res[k][1][300]) <= GRAD_RTOL * gscale
    _say("[gpu] 512^2: mixed-precision primal + adjoint")
    with mgfvm.Batch([pairs[k] for k in idx], precision=mgfvm.PREC_MIXED) as bm:
        ncm, cvm = bm.solve()
        assert all(cvm)
        resm = gadj.batch_adjoint(bm, 0)
    for q, k in enumerate(idx):
        _close(resm[q][1], res[k][1], 1e-5, f"mixed inlet {k}")
        _close(resm[q][2], res[k][2], 1e-5, f"mixed amuf {k}")


def test_batched_adjoint_mixed_shapes_equals_singles(tmp_path):
    """A batch whose problems differ in size, depth and obstacle level (32^2
    with 3 levels twice, at two viscosities, and 64 x 32 with obstacles
    coarsened below their level): every member's batched gradients of both
    objectives equal those of the same problem solved and differentiated
    alone, to 1e-12."""
    import mgfvm
    from mgfvm import adjoint as gadj
    (p1, f1), (p3, f3) = _case32(), _case64_coarse_obstacles()
    cases = [(p1, f1), (p1.replace(amuf=0.8 * p1.amuf), f1), (p3, f3)]
    pairs = [p.write(str(tmp_path / f"m{k}")) for k, (p, _) in enumerate(cases)]
    _say("[gpu] batched adjoint over mixed shapes vs single problems")
    with mgfvm.Batch(pairs) as bt:
        for b, (_, prof) in enumerate(cases):
            bt.set_inlet_profile(b, prof)
        nc, cv = bt.solve()
        assert all(cv)
        res = {obj: gadj.batch_adjoint(bt, obj) for obj in (0, 1)}
    for b, ((pin, ppat), (_, prof)) in enumerate(zip(pairs, cases)):
        with mgfvm.Solver(pin, ppat, inlet_u=prof) as s:
            assert s.solve()[2]
            for obj in (0, 1):
                gi, ga, gp = gadj.adjoint(s, obj)
                _close(res[obj][b][1], gi, 1e-12, f"member {b} obj {obj} inlet rows")
                _close(res[obj][b][2], ga, 1e-12, f"member {b} obj {obj} amuf")
                _close(res[obj][b][3], gp, 1e-12, f"member {b} obj {obj} prsc")


def test_blocked_sweeps_equal_single_colour(tmp_path):
    """The preconditioner's sweeps run several per launch on LDS tiles (two
    launches per field, to a scratch array and back); the same sweeps as
    one launch per colour give the same bits: every gradient, J, iteration
    count and final residual, for both objectives of the 256^2
    differentiation workload and for a batch of two 512^2 problems."""
    import mgfvm
    from mgfvm import adjoint as gadj
    from mgfvm.inputs import adjoint_batch, sensitivity_256
    p, inlet = sensitivity_256(1)
    pin, ppat = p.write(str(tmp_path / "g256"))
    probs = [q.replace(convergence=1e-10) for q in adjoint_batch(2)]
    pairs = [q.write(str(tmp_path / f"b{k}")) for k, q in enumerate(probs)]
    _say("[gpu] blocked vs single-colour adjoint sweeps: 256^2 and 2 x 512^2")
    with mgfvm.Solver(pin, ppat, inlet_u=np.array(inlet)) as s, mgfvm.Batch(pairs) as bt:
        assert s.solve()[2]
        bt.solve()
        for solver, objs in ((s, (0, 1)), (bt, (0,))):
            for obj in objs:
                a = gadj.batch_adjoint(solver, obj, gadj.opts(), allow_unconverged=True)
                b = gadj.batch_adjoint(solver, obj, gadj.opts(single_colour=True),
                                       allow_unconverged=True)
                for ra, rb in zip(a, b):
                    assert ra[4] == rb[4]
                    va = np.concatenate([[ra[0], ra[2], ra[3], ra[5]], ra[1]])
                    vb = np.concatenate([[rb[0], rb[2], rb[3], rb[5]], rb[1]])
                    assert np.array_equal(va.view(np.uint64), vb.view(np.uint64)), obj


def test_unconverged_solver_is_refused(tmp_path):
    import mgfvm
    from mgfvm import _lib
    from mgfvm import adjoint as gadj
    prob, prof = _case32()
    pin, ppat = prob.replace(nvcycle=3).write(str(tmp_path / "u"))
    with mgfvm.Solver(pin, ppat, inlet_u=prof) as s:
        s.solve()
        with pytest.raises(_lib.MgfError) as e:
            gadj.adjoint(s, 0)
        assert e.value.code == -4
        with pytest.raises(_lib.MgfError):
            gadj.adjoint(s, 2)


def test_discrete_inlets_and_outlets_match_forward_mode(tmp_path):
    """Discrete inlets and outlets (on level 0) set fixed ghost values, which
    the adjoint's rows model: on the 32^2 case with an inlet (its sc 1) and
    an outlet over part of the east face, every direction of both objectives
    equals the forward-mode library's."""
    import mgfvm
    from mgfvm import adjoint as gadj
    from mgfvm.tangent import gradient_c
    prob, prof = _case32()
    um = float(prof.mean())
    prob = prob.replace(ig_inlt=1, ig_outlt=1, inlets=[(1, 32, 1.0, 0.0, 1.0)],
                        outlets=[(9, 24, um, 0.0, 0.0)])
    pin, ppat = prob.write(str(tmp_path / "io"))
    _say("[gpu] discrete inlets/outlets: forward mode vs adjoint")
    J, gin, gmu, gpr, conv = gradient_c(pin, ppat, inlet_u=prof)
    assert conv
    with mgfvm.Solver(pin, ppat, inlet_u=prof) as s:
        assert s.solve()[2]
        for obj in (0, 1):
            gi, ga, gp = gadj.adjoint(s, obj)
            _close(gi, gin[obj], GRAD_RTOL, f"obj {obj} inlet rows")
            _close(ga, gmu[obj], GRAD_RTOL, f"obj {obj} amuf")
            if obj == 1:
                _close(gp, gpr[obj], GRAD_RTOL, "obj 1 prsc")


def test_unmodelled_boundaries_are_refused(tmp_path):
    """The adjoint's rows hold every ghost value fixed (ibc 1): a problem
    with a p' Dirichlet row, periodic or outflow side is refused with
    MGF_ERR_CONFIG rather than given a wrong gradient; its objectives can
    still be read."""
    import mgfvm
    from mgfvm import _lib
    from mgfvm import adjoint as gadj
    prob, prof = _case32()
    for k, ibc in enumerate(((2, 2, 2, 2), (3, 1, 3, 1), (1, 4, 1, 1))):
        pin, ppat = prob.replace(ibc=ibc, nvcycle=40, convergence=0.0).write(
            str(tmp_path / f"b{k}"))
        with mgfvm.Solver(pin, ppat, inlet_u=prof) as s:
            s.solve()
            gadj.objectives(s)
            with pytest.raises(_lib.MgfError) as e:
                gadj.batch_adjoint(s, 0, gadj.opts(only_converged=False))
            assert e.value.code == -4 and "ibc 1" in str(e.value), ibc


def test_torch_function_backward_is_the_adjoint(tmp_path):
    """torch.autograd.Function: backward returns the adjoint gradient for
    tensor inputs (on the inputs' device and dtype), None for float
    hyperparameters, and matches a directional central difference of the
    forward."""
    import torch
    from mgfvm import adjoint as gadj
    prob, prof = _case32()
    F = gadj.torch_function()
    dev = "cuda:0"
    inlet = torch.tensor(prof, dtype=torch.float64, device=dev, requires_grad=True)
    amuf = torch.tensor(prob.amuf, dtype=torch.float64, device=dev, requires_grad=True)
    _say("[gpu] torch Function forward/backward")
    J = F.apply(inlet, amuf, prob.prsc, prob, 0)
    J.backward()
    assert inlet.grad is not None and inlet.grad.device == inlet.device
    assert amuf.grad is not None and amuf.grad.dtype == torch.float64
    # float prsc: no gradient slot, no error
    d = torch.zeros_like(inlet)
    d[3] = 1.0
    d[20] = -0.5
    h = 1e-5
    with torch.no_grad():
        Jp = F.apply(inlet + h * d, amuf, prob.prsc, prob, 0)
        Jm = F.apply(inlet - h * d, amuf, prob.prsc, prob, 0)
    fd = float((Jp - Jm) / (2 * h))
    # the profile's mean moves the east u too (mass-consistent), as the
    # gradient is defined
    an = float((inlet.grad * d).sum())
    assert abs(an - fd) <= 1e-5 * max(abs(fd), 1e-12), (an, fd)
