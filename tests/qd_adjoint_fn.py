"""torch.autograd.Function over the GPU discrete adjoint of the QD step
(verification harness only: the product path is the C ABI).

QdMisfit.apply(theta_a, theta_b, theta_Dc, ctx) returns the misfit J of the
mediator in ctx with a, b, Dc = P theta (P the hat-function matrix of the
profiles); its backward is P^T times the adjoint's per-node gradient.
"""
import numpy as np
import torch


class MisfitContext:
    def __init__(self, mediator, P, dts, stations, data, **adjoint_kw):
        self.m, self.P, self.dts, self.stations, self.data = mediator, P, dts, stations, data
        self.kw = adjoint_kw


class QdMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta_a, theta_b, theta_dc, mc):
        P = mc.P
        f = [P @ t.detach().cpu().numpy().astype(np.float64) for t in (theta_a, theta_b, theta_dc)]
        mc.m.set_friction(*f)
        ga, gb, gdc, st = mc.m.adjoint(mc.dts, mc.stations, mc.data, **mc.kw)
        ctx.save_for_backward(*[torch.from_numpy(P.T @ g) for g in (ga, gb, gdc)])
        return torch.tensor(st["J"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        ga, gb, gdc = ctx.saved_tensors
        return gJ * ga, gJ * gb, gJ * gdc, None
