"""The robust seam energy as a torch.autograd.Function (PyTorch-ROCm).

HuberEnergy.apply(xi, batch, delta, precision) returns the per-pose energies E (k,) of a
HuberBatch at the twists xi (k, 6), computed by the device kernel; its backward launches the
hand-written gradient kernel (dE/dxi = (J_l(omega)^T sum psi (R T0 s) x n, sum psi n)) and scales
it by the incoming gradient. The twists may live on any device; values cross to the library as
f64 host arrays (6 doubles per pose), the point data stays on the GPU.
"""

import torch


class HuberEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, batch, delta, precision=64):
        x = xi.detach().to("cpu", torch.float64).contiguous().numpy()
        E = batch.energy(x, delta, precision)
        ctx.batch, ctx.delta, ctx.precision = batch, float(delta), int(precision)
        ctx.save_for_backward(xi)
        return torch.from_numpy(E).to(device=xi.device, dtype=xi.dtype)

    @staticmethod
    def backward(ctx, grad_e):
        (xi,) = ctx.saved_tensors
        x = xi.detach().to("cpu", torch.float64).contiguous().numpy()
        _, G = ctx.batch.energy(x, ctx.delta, ctx.precision, grad=True)
        g = torch.from_numpy(G).to(device=xi.device, dtype=xi.dtype)
        return grad_e[:, None] * g, None, None, None


def huber_energy(xi, batch, delta, precision=64):
    return HuberEnergy.apply(xi, batch, delta, precision)
