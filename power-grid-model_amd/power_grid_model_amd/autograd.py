"""Differentiable power-flow objective for PyTorch.

PowerFlowLoss.apply(load_p, load_q, line_r1, line_x1, input_data) returns, for every scenario s,

    L_s = sum over energised nodes (u_pu - 1)^2 + sum over lines (p_from + p_to) [MW]

of a symmetric Newton-Raphson batch whose scenario s sets every sym_load's p_specified / q_specified to
row s of load_p / load_q [W, var], on the model input_data with every line's r1 / x1 replaced by line_r1 /
line_x1 [ohm]. The forward pass solves the batch on the GPU; the backward pass is the adjoint
(J^T lambda = dL/dx with the Jacobian factorised again, in FP64, at the converged state), so one backward
costs about one extra Newton iteration. The line parameters are shared by all scenarios: their gradient
is summed over the scenarios, weighted by the incoming gradient of each L_s.

The batch is built with the adjoint layout (every topology group a plain group), so scenarios may differ
in their load statuses as well as their powers. Scenarios that fail to converge give NaN losses and zero
gradients.
"""

from __future__ import annotations

import numpy as np
import torch

from .meta import initialize_array
from .model import DeviceBatch, PowerGridModel

OUTPUT = {"node": ["u_pu", "energized"], "line": ["p_from", "p_to"]}


def _solve(input_data, load_p, load_q, line_r1, line_x1, factor_precision, error_tolerance, max_iterations):
    inp = {k: v.copy() for k, v in input_data.items()}
    inp["line"]["r1"] = line_r1
    inp["line"]["x1"] = line_x1
    loads = inp["sym_load"]
    n_scn = load_p.shape[0]
    upd = initialize_array("update", "sym_load", (n_scn, len(loads)))
    upd["id"] = loads["id"][None, :]
    upd["p_specified"] = load_p
    upd["q_specified"] = load_q
    batch = DeviceBatch(PowerGridModel(inp), {"sym_load": upd}, factor_precision=factor_precision,
                        error_tolerance=error_tolerance, max_iterations=max_iterations, adjoint_layout=True)
    batch.set_output(OUTPUT)
    batch.run()
    out = batch.fetch()
    u, en = out["node"]["u_pu"], out["node"]["energized"]
    loss = np.sum(np.where(en == 1, (u - 1.0) ** 2, 0.0), axis=1)
    loss = loss + np.sum(out["line"]["p_from"] + out["line"]["p_to"], axis=1) / 1e6
    loss[batch.iterations() == 0] = np.nan
    return loss, batch


class PowerFlowLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, load_p, load_q, line_r1, line_x1, input_data, factor_precision="fp64", error_tolerance=1e-8,
                max_iterations=20):
        args = [t.detach().cpu().to(torch.float64).numpy() for t in (load_p, load_q, line_r1, line_x1)]
        loss, batch = _solve(input_data, *args, factor_precision, error_tolerance, max_iterations)
        # the adjoint runs in backward only (a forward without backward costs one solve); the batch keeps the
        # converged state on the device until then
        ctx.batch = batch
        ctx.inputs = [(t.dtype, t.device) for t in (load_p, load_q, line_r1, line_x1)]
        return torch.from_numpy(loss).to(device=load_p.device, dtype=load_p.dtype)

    @staticmethod
    def backward(ctx, grad_loss):
        grads = ctx.batch.adjoint()
        ctx.batch = None
        lg, ln = grads["sym_load"], grads["line"]
        gp, gq, gr, gx = (torch.from_numpy(np.ascontiguousarray(a)) for a in
                          (lg["p_specified"], lg["q_specified"], ln["r1"], ln["x1"]))
        w = grad_loss.detach().cpu().to(torch.float64)[:, None]
        out = [w * gp, w * gq, (w * gr).sum(0), (w * gx).sum(0)]
        return tuple(o.to(device=dev, dtype=dt) for o, (dt, dev) in zip(out, ctx.inputs)) + (None, None, None, None)
