"""Compliance and its design gradient as a differentiable op (libtopgpu tg_compliance_grad).

    op = ComplianceOp(model, rMin=3.0)            # PDE filter radius (None: no filter)
    x = torch.full((Ne,), 0.12, dtype=torch.float64, device='cuda', requires_grad=True)
    c = compliance(x, op)                         # torch.autograd.Function
    c.backward()                                  # x.grad = dc/dx

c = F'U with K(E(filter(x))) U = F, E = Emin + xPhys^p (E0 - Emin); the gradient is the SIMP
sensitivity -p (E0 - Emin) xPhys^(p-1) u_e'K0 u_e (TOP_GPU.m:158-174, 189-209, without the
1/cSolid normalisation of the reference's objective) chained back through the symmetric PDE filter
(TOP_GPU.m:214-218).  Compliance is self-adjoint, so the backward pass reuses the forward solve.
The solve warm-starts from the previous call's displacement, like TOP_GPU.m:186.
"""
import ctypes

import torch

from ._lib import check, lib
from .solver import Solver
from .top_gpu import PDEFilter


class ComplianceOp:
    def __init__(self, model, rMin=None, Emin=1e-6, E0=1.0, penal=3.0, tol=1e-3, max_it=800, filter_tol=1e-4,
                 solver=None):
        self.model = model
        self.solver = solver if solver is not None else Solver(model)
        self.filter = PDEFilter(model, rMin) if rMin else None
        self.Emin, self.E0, self.penal = Emin, E0, penal
        self.tol, self.max_it, self.filter_tol = tol, max_it, filter_tol
        self.U = torch.zeros(self.solver.ndof(1), dtype=torch.float64, device='cuda')
        self.warm = False
        self.last_its = 0

    def __call__(self, x, grad=True):
        """(c, dc/dx or None); also keeps the displacement in self.U."""
        x = x.detach().to(torch.float64).contiguous()
        g = torch.empty_like(x) if grad else None
        c = ctypes.c_double()
        its = ctypes.c_int32()
        check(lib().tg_compliance_grad(self.solver.h, self.filter.h if self.filter else None,
                                       ctypes.c_void_p(x.data_ptr()), self.Emin, self.E0, self.penal, self.tol,
                                       int(self.max_it), self.filter_tol, int(self.warm),
                                       ctypes.c_void_p(self.U.data_ptr()), ctypes.byref(c),
                                       ctypes.c_void_p(g.data_ptr()) if grad else None, ctypes.byref(its)))
        self.warm = True
        self.last_its = its.value
        return c.value, g


class Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op):
        c, g = op(x, grad=True)
        ctx.save_for_backward(g)
        return torch.tensor(c, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


def compliance(x, op):
    return Compliance.apply(x, op)
