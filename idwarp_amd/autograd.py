"""torch.autograd wrapper of the device volume warp.

``warp(mesh, xs)`` maps a flat float64 CUDA tensor of surface coordinates
(3 * nSurf, the ordering of USMesh.getSurfaceCoordinates) to the warped
volume grid (3 * nVol, warp ordering).  Its backward is the reverse product
dXs = (dXv/dXs)^T dXv of warpDeriv (src/warp/warpDeriv.F90), computed by
the library's reverse kernels, so ``xv.backward(g)`` fills ``xs.grad`` with
what USMesh.warpDeriv(g) / getdXs() returns.  Forward-mode AD (torch.func.jvp,
torch.autograd.forward_ad) uses the tangent of warpDerivFwd.F90.  All
directions run on torch's current stream through idw_warp_mesh_dev,
idw_warp_deriv_dev and idw_warp_deriv_fwd_dev; torch only provides the
device memory and the stream.
"""
import ctypes

import torch

from . import _lib


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_surface(mesh, xs):
    if not isinstance(xs, torch.Tensor) or not xs.is_cuda or xs.dtype != torch.float64:
        raise TypeError("xs must be a float64 CUDA tensor")
    if xs.numel() != 3 * mesh.nSurf:
        raise ValueError("xs has %d values, the surface has %d" % (xs.numel(), 3 * mesh.nSurf))


class WarpFunction(torch.autograd.Function):
    """Xv = warpMesh(Xs); backward: dXs = warpDeriv(dXv); jvp: dXv =
    warpDerivFwd(dXs)."""

    @staticmethod
    def forward(ctx, xs, mesh):
        _check_surface(mesh, xs)
        xs = xs.contiguous()
        mesh._setMeshOptions()
        xv = torch.empty(3 * mesh.nVol, dtype=torch.float64, device=xs.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(xs.device).cuda_stream)
        _lib.check(mesh._L.idw_warp_mesh_dev(mesh.ctx, _vp(xs), _vp(xv), stream))
        ctx.mesh = mesh
        ctx.save_for_backward(xs)
        ctx.save_for_forward(xs)
        return xv

    @staticmethod
    def jvp(ctx, dxs, _mesh_tangent):
        (xs,) = ctx.saved_tensors
        mesh = ctx.mesh
        dxs = dxs.to(torch.float64).contiguous()
        dxv = torch.empty(3 * mesh.nVol, dtype=torch.float64, device=xs.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(xs.device).cuda_stream)
        _lib.check(mesh._L.idw_warp_deriv_fwd_dev(mesh.ctx, _vp(xs), _vp(dxs), _vp(dxv), stream))
        return dxv

    @staticmethod
    def backward(ctx, gxv):
        (xs,) = ctx.saved_tensors
        mesh = ctx.mesh
        gxv = gxv.to(torch.float64).contiguous()
        gxs = torch.empty_like(xs)
        stream = ctypes.c_void_p(torch.cuda.current_stream(xs.device).cuda_stream)
        _lib.check(mesh._L.idw_warp_deriv_dev(mesh.ctx, _vp(xs), _vp(gxv), _vp(gxs), stream))
        return gxs, None


def warp(mesh, xs):
    """Differentiable warp of ``mesh`` (an idwarp_amd.USMesh with its surface
    definition set) for the surface coordinates ``xs``."""
    return WarpFunction.apply(xs, mesh)
