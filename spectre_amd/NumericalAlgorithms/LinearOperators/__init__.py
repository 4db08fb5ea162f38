"""partial_derivative and logical_partial_derivative with the names and
argument order of spectre.NumericalAlgorithms.LinearOperators
(NumericalAlgorithms/LinearOperators/Python/PartialDerivatives.cpp:20-40),
on torch tensors on the HIP device, differentiable with torch.autograd.

* tensor: float64 [nc, Np] (the components of a tensor of DataVectors, as
  np.array(tensor) gives them in the reference) or a batch [E, nc, Np];
* mesh: Mesh[3](extents, Basis.Legendre, Quadrature.GaussLobatto);
* inv_jacobian: [9, Np] (or [E, 9, Np]), block j + 3 i = d xi^j / d x^i.

partial_derivative returns [3 nc, Np] with the derivative index fastest
(block 3 c + i = d_i of component c), as the reference prepends the spatial
index; logical_partial_derivative likewise with logical derivatives.

Gradients: with respect to the tensor, the transposed derivative kernel
(sum_j D_j^T applied to the inverse-Jacobian-weighted adjoint); with respect
to the inverse Jacobian, sum_c grad_{3c+i} * logical_{3c+j} from the logical
derivatives.  All device work goes through the pybind11 module
spectre_amd._Pybindings and the C ABI; there is no CPU path.
"""
import torch
from torch.autograd.function import once_differentiable

from ... import _Pybindings as _pyb


def _identity_inverse_jacobian(like, npts, batch):
    eye = torch.zeros((9, npts), dtype=torch.float64, device=like.device)
    eye[0] = eye[4] = eye[8] = 1.0
    return eye.expand(batch, 9, npts).contiguous() if batch is not None else eye


class PartialDerivative(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tensor, mesh, inv_jacobian):
        t = tensor.detach().contiguous()
        inv = inv_jacobian.detach().contiguous()
        ctx.mesh = mesh
        ctx.save_for_backward(t, inv)
        return _pyb.partial_derivative(t, mesh, inv)

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        t, inv = ctx.saved_tensors
        grad = grad.contiguous()
        g_tensor = g_inv = None
        if ctx.needs_input_grad[0]:
            g_tensor = _pyb.partial_derivative_adjoint(grad, ctx.mesh, inv)
        if ctx.needs_input_grad[2]:
            logical = _pyb.logical_partial_derivative(t, ctx.mesh)
            lead = grad.shape[:-2]
            nc, npts = t.shape[-2], t.shape[-1]
            gv = grad.reshape(*lead, nc, 3, 1, npts)      # [.., c, i, 1, p]
            lv = logical.reshape(*lead, nc, 1, 3, npts)   # [.., c, 1, j, p]
            g_inv = (gv * lv).sum(dim=-4).reshape(*lead, 9, npts)  # block 3 i + j
        return g_tensor, None, g_inv


class LogicalPartialDerivative(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tensor, mesh):
        t = tensor.detach().contiguous()
        ctx.mesh = mesh
        ctx.batch = t.shape[0] if t.dim() == 3 else None
        return _pyb.logical_partial_derivative(t, mesh)

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        grad = grad.contiguous()
        eye = _identity_inverse_jacobian(grad, grad.shape[-1], ctx.batch)
        return _pyb.partial_derivative_adjoint(grad, ctx.mesh, eye), None


def partial_derivative(tensor, mesh, inv_jacobian):
    """Inertial partial derivatives of every component of `tensor`."""
    return PartialDerivative.apply(tensor, mesh, inv_jacobian)


def logical_partial_derivative(tensor, mesh):
    """Element-logical partial derivatives of every component of `tensor`."""
    return LogicalPartialDerivative.apply(tensor, mesh)


__all__ = ["partial_derivative", "logical_partial_derivative", "PartialDerivative",
           "LogicalPartialDerivative"]
