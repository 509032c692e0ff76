"""The compliance-gradient op (tg_compliance_grad, adjoint.Compliance) against central finite
differences and the reference's dc formula.

Finite differences of c(x) = F'K(E(filter(x)))^-1 F at h = 1e-6 cannot be formed as
c(x+h) - c(x-h) in FP64 at these sizes: the difference is ~1e-12 of c.  The compliance differences are
therefore evaluated with the exact identity c(E') - c(E) = -U' (K(E') - K(E)) U = -sum_e dE_e u_e'K0 u'_e
(U, U' the two solutions), which has no cancellation; dE follows from the linear filter,
filter(x + s h e_i) - filter(x) = filter(s h e_i), and u_e'K0 u'_e from the polarisation
(ce(U+U') - ce(U-U')) / 4.  Solves and filters run to relative residual 1e-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P, EMIN, E0 = 3.0, 1e-6, 1.0


def _fd_check(shape, precision, n_vox, seed, rMin):
    import torch
    from top_gpu_on_matlab_amd import _lib
    from top_gpu_on_matlab_amd import top_gpu as G
    from top_gpu_on_matlab_amd.adjoint import ComplianceOp, compliance
    from top_gpu_on_matlab_amd.model import cuboid_model
    model = cuboid_model(*shape, 1, precision=precision)
    Ne = model.num_elements
    rng = np.random.default_rng(12345)
    x = torch.from_numpy(0.12 + rng.uniform(-0.05, 0.05, Ne)).cuda()
    op = ComplianceOp(model, rMin=rMin, Emin=EMIN, E0=E0, penal=P, tol=1e-12, max_it=2000, filter_tol=1e-12)
    filt = G.PDEFilter(model, rMin, tol=1e-13, max_it=4000)
    # forward + backward through the autograd op
    xr = x.clone().requires_grad_(True)
    c = compliance(xr, op)
    c.backward()
    g = xr.grad.clone()
    U0 = op.U.clone()
    c0 = float(c.detach())
    # the reference's normalised sensitivity, filtered: dc_ref = filter(-p (E0-Emin) xPhys^(p-1) ce / cSolid)
    xPhys = filt(x)
    ce = G.unit_compliance(op.solver, U0)
    dc_ref = filt(G.sensitivity(xPhys, ce, 2.0, EMIN, E0, P))
    assert torch.allclose(dc_ref * 2.0, g, rtol=1e-9, atol=1e-12 * float(g.abs().max()))
    assert abs(float(torch.dot(G.simp(xPhys, EMIN, E0, P), ce)) - c0) < 1e-10 * c0   # c = sum E_e ce_e = F'U

    vox = np.random.default_rng(seed).choice(Ne, n_vox, replace=False)
    h = 1e-6
    errs = []
    for i in vox:
        dcs = []
        for s in (1.0, -1.0):
            xs = x.clone()
            xs[i] += s * h
            op.warm = True
            op.U.copy_(U0)
            op(xs, grad=False)
            Us = op.U.clone()
            e_i = torch.zeros_like(x)
            e_i[i] = s * h
            dxp = filt(e_i)
            dE = (E0 - EMIN) * ((xPhys + dxp) ** 3 - xPhys ** 3)
            w = (G.unit_compliance(op.solver, U0 + Us) - G.unit_compliance(op.solver, U0 - Us)) / 4
            dcs.append(-float(torch.dot(dE, w)))
        fd = (dcs[0] - dcs[1]) / (2 * h)
        errs.append(abs(fd - float(g[i])) / abs(float(g[i])))
    print('\n[%s %s] c = %.10e; FD vs gradient on %d voxels: max rel %.2e, median %.2e'
          % (shape, precision, c0, n_vox, max(errs), float(np.median(errs))), flush=True)
    op.U.copy_(U0)
    return max(errs)


def test_gradient_matches_central_differences_small():
    assert _fd_check((16, 32, 16), 'fp64', 12, seed=3, rMin=2.0) < 1e-5


def test_gradient_matches_central_differences_128x256x128_mixed():
    """The adjoint run: 128 x 256 x 128 cantilever, ft = 2, rMin = 3, x0 = 0.12 + U(-0.05, 0.05),
    mixed-precision forward (FP32 V-cycle inside FP64 PCG), 64 seeded voxels."""
    assert _fd_check((128, 256, 128), 'mixed', 64, seed=777, rMin=3.0) < 1e-5
