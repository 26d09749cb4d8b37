"""PyTorch face of the C ABI.

`residual(u, system)` evaluates the SCF residual f(u) of an `sfx.System` on
device tensors and is a torch.autograd.Function whose forward-mode derivative
is the hand-written tangent propagator sweep (`sfx_jvp`): under
`torch.autograd.forward_ad` (or `torch.func.jvp`) the tangent of f is the exact
Jacobian-vector product J(u) v, not a finite difference. Its reverse-mode
derivative is the transposed product J(u)^T w (`sfx_vjp`), so `backward()` and
`torch.autograd.grad` work as well. The iteration vector uses the layout of
`sfx_residual` (the reference's SF_Solve / MPS layout).
"""
import torch

from . import sfx

_PREC = ("fp64", "fp32")


class _Residual(torch.autograd.Function):
    @staticmethod
    def forward(u, system, precision):
        return system.residual(u.detach().contiguous(), precision=precision)

    @staticmethod
    def setup_context(ctx, inputs, output):
        u, system, precision = inputs
        ctx.system, ctx.precision = system, precision
        ctx.u = u.detach().contiguous()

    @staticmethod
    def jvp(ctx, du, _dsystem, _dprecision):
        return ctx.system.jvp(ctx.u, du.detach().contiguous(), precision=ctx.precision)

    @staticmethod
    def backward(ctx, grad):
        return ctx.system.vjp(ctx.u, grad.detach().contiguous(), precision=ctx.precision), None, None


def residual(u: torch.Tensor, system: "sfx.System", precision: str = "fp64") -> torch.Tensor:
    """f(u) on the GPU; u is a float64 CUDA tensor of system.nvar values."""
    if precision not in _PREC:
        raise ValueError(f"precision must be one of {_PREC}")
    if not (isinstance(u, torch.Tensor) and u.is_cuda and u.dtype == torch.float64):
        raise TypeError("u must be a float64 CUDA tensor")
    if u.numel() != system.nvar:
        raise ValueError(f"u has {u.numel()} values, the system {system.nvar}")
    return _Residual.apply(u, system, precision)


def jvp(u: torch.Tensor, v: torch.Tensor, system: "sfx.System", precision: str = "fp64"):
    """(f(u), J(u) v) through torch's forward-mode AD."""
    import torch.autograd.forward_ad as fwAD

    with fwAD.dual_level():
        f = residual(fwAD.make_dual(u, v), system, precision)
        out = fwAD.unpack_dual(f)
        return out.primal, out.tangent
