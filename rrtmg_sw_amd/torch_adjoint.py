"""torch.autograd.Function over the fused McICA shortwave path.

    toa_up, swhr = McicaShortwave.apply(tlay, h2ovmr, o3vmr, clwp, ciwp, alb, cols, dev)

Inputs are device float64 tensors in the library's layout (torch shape = reversed Fortran
shape): tlay, h2ovmr, o3vmr, clwp, ciwp are (nlay, ncol) = Fortran (ncol, nlay); alb is
(4, ncol): asdir, asdif, aldir, aldif. `cols` (a workloads.Columns) supplies the sizes and flags,
`dev` the other device inputs (api.device_inputs). Outputs: the total-sky upward flux at the top
level (ncol,) and the heating rates (nlay, ncol). The forward runs the library's fused forward
with exact exponentials (the function whose derivative the adjoint computes); backward calls
the hand-written adjoint (rrtmgsw_mcica_sw_adjoint_dev). McICA masks are held fixed.
"""
import torch

from . import api

_NAMES = ("tlay", "h2ovmr", "o3vmr", "clwp", "ciwp")
_ALB = ("asdir", "asdif", "aldir", "aldif")


def _inputs(dev, xs, alb):
    d = dict(dev)
    for k, x in zip(_NAMES, xs):
        d[k] = x.detach().contiguous()
    for i, k in enumerate(_ALB):
        d[k] = alb[i].detach().contiguous()
    return d


class McicaShortwave(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tlay, h2ovmr, o3vmr, clwp, ciwp, alb, cols, dev, precision="exact"):
        # the adjoint is the derivative of the exact-exponential forward only: the table forward
        # is piecewise constant in the optical depths, the mixed form another function again
        if precision != "exact":
            raise ValueError("McicaShortwave differentiates the exact-exponential forward: "
                             "precision must be 'exact'")
        d = _inputs(dev, (tlay, h2ovmr, o3vmr, clwp, ciwp), alb)
        out = api.mcica_rrtmg_sw(cols, d, precision=precision)
        ctx.cols, ctx.d = cols, d
        return out["swuflx"][cols.nlay].clone(), out["swhr"]

    @staticmethod
    def backward(ctx, g_toa, g_hr):
        cols = ctx.cols
        if g_toa is None:
            g_toa = torch.zeros(cols.ncol, dtype=torch.float64, device="cuda")
        if g_hr is None:
            g_hr = torch.zeros(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
        g = api.mcica_rrtmg_sw_adjoint(cols, ctx.d, g_toa.double().contiguous(),
                                       g_hr.double().contiguous())
        return (g["tlay"], g["h2ovmr"], g["o3vmr"], g["clwp"], g["ciwp"], g["alb"], None, None,
                None)
