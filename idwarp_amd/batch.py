"""Graph-recorded batches of warps and reverse products (idw_batch_* in
include/idwarp_hip.h).

    b = Batch()
    for xs, xv in zip(surfaces, outputs):
        b.add_warp(mesh, xs, xv)
    b.add_warp_deriv(mesh, xs0, dxv, dxs)
    b.run()            # first run: executes, records and picks a graph layout
    b.run()            # later runs: one graph launch

All tensors are float64 CUDA tensors on the batch's device, flat and
contiguous: surfaces 3*nSurf, volume fields 3*nVol.  The batch keeps
references to them, so they stay alive as long as it does.
"""
import ctypes

import torch

from . import _lib


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class Batch:
    def __init__(self, device=0):
        self._L = _lib.load()
        self._h = ctypes.c_void_p(0)
        _lib.check(self._L.idw_batch_create(ctypes.byref(self._h), int(device)))
        self.device = device
        self._keep = []
        self._meshes = []

    def _check(self, mesh, *tensors):
        for t in tensors:
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise TypeError("batch operands must be contiguous float64 CUDA tensors")
        self._keep.extend(tensors)
        if mesh not in self._meshes:
            self._meshes.append(mesh)
        mesh._setMeshOptions()

    def add_warp(self, mesh, xs, xv):
        """Xv = warpMesh(Xs) for the surface xs (3*nSurf) into xv (3*nVol)."""
        self._check(mesh, xs, xv)
        if xs.numel() != 3 * mesh.nSurf or xv.numel() != 3 * mesh.nVol:
            raise ValueError("xs must hold 3*nSurf and xv 3*nVol values")
        _lib.check(self._L.idw_batch_add_warp(self._h, mesh.ctx, _vp(xs), _vp(xv)))

    def add_warp_deriv(self, mesh, xs, dxv, dxs):
        """dxs = (dXv/dXs)^T dxv at the surface xs."""
        self._check(mesh, xs, dxv, dxs)
        if xs.numel() != 3 * mesh.nSurf or dxs.numel() != 3 * mesh.nSurf or dxv.numel() != 3 * mesh.nVol:
            raise ValueError("xs and dxs must hold 3*nSurf, dxv 3*nVol values")
        _lib.check(self._L.idw_batch_add_warp_deriv(self._h, mesh.ctx, _vp(xs), _vp(dxv), _vp(dxs)))

    def run(self, stream=None):
        """Enqueue the batch on torch's current stream (or ``stream``)."""
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        _lib.check(self._L.idw_batch_run(self._h, ctypes.c_void_p(s.cuda_stream)))

    def info(self):
        """(layout, [ms of one replay as branches, as a chain]) of the recording."""
        layout = ctypes.c_int(-1)
        ms = (ctypes.c_double * 2)()
        _lib.check(self._L.idw_batch_info(self._h, ctypes.byref(layout), ms))
        return layout.value, [ms[0], ms[1]]

    def stats(self):
        """(recordings, graph launches) so far: a recording is the eager
        warm-up run, the timed layouts and the graph kept; later runs that
        find every mesh unchanged only launch the graph."""
        rec, lau = ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.check(self._L.idw_batch_stats(self._h, ctypes.byref(rec), ctypes.byref(lau)))
        return rec.value, lau.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.idw_batch_destroy(h)
            self._h = None
