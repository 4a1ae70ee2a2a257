"""torch.autograd.Function around the library's network forward and hand-written backward, for
the PyTorch-ROCm test harness only (the library itself takes no torch types).

NetFunction maps a vector of selected parameters of one net (the rest held fixed) to the net's
outputs on a batch: forward = nrrs_params_set + nrrs_statnet_infer / nrrs_rrsnet_infer, backward =
nrrs_backward with the incoming dL/dy (NRRS_LOSS_EXTERNAL) + nrrs_grads_get. torch then composes
it with any loss written in torch, and torch.autograd.gradcheck can test it by central
differences."""
import numpy as np
import torch

import nrrs_amd as na


class NetFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sel_values, nctx, net, vb, film, stats, base_params, sel_idx, loss_scale):
        p = base_params.copy()
        p[sel_idx] = sel_values.detach().cpu().numpy().astype(np.float32)
        nctx.params_set(net, p)
        if net == na.STATNET:
            y = nctx.statnet_infer(vb)  # [6][n]
        else:
            y = nctx.rrsnet_infer(vb, film, stats)[None, :]  # [1][n]
        fctx.nrrs = (nctx, net, vb, film, stats, sel_idx, loss_scale, sel_values.device)
        return y.to(sel_values.dtype)

    @staticmethod
    def backward(fctx, dy):
        nctx, net, vb, film, stats, sel_idx, loss_scale, device = fctx.nrrs
        nctx.grads_zero(net)
        d = dy.to(torch.float32).contiguous()
        na.check(nctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, stats=stats, dy=d, loss_scale=loss_scale),
                 "backward")
        g = nctx.grads_get(net) / loss_scale
        grad = torch.from_numpy(g[sel_idx].astype(np.float64)).to(device)
        return grad, None, None, None, None, None, None, None, None


def net_outputs(sel_values, nctx, net, vb, film, stats, base_params, sel_idx, loss_scale=1.0):
    return NetFunction.apply(sel_values, nctx, net, vb, film, stats, base_params, sel_idx, loss_scale)
