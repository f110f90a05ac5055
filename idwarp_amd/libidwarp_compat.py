"""A ``libidwarp``-shaped object over the MI355X library.

The reference's USMesh talks to its f2py module through ``self.warp``
(idwarp/UnstructuredMesh.py:96, the module of src/f2py/libidwarp.pyf).
``LibIdwarp`` offers the same subroutines, argument shapes and module
variables for the volume-warp path, so the reference class can run on this
library by replacing that one assignment:

    self.warp = idwarp_amd.libidwarp_compat.LibIdwarp()

Option values are read from ``gridinput`` / ``kd_tree`` when a warp or a
reverse product starts, as the Fortran reads its module variables then.
Grid file readers (readcgns, readplot3d, ...) are not part of the path:
the volume grid comes in through ``createcommongrid``, as the reference's
OpenFOAM reader passes it (UnstructuredMesh.py:1004-1049).
"""
import ctypes
import types

import numpy as np

from . import _lib


class _Module(types.SimpleNamespace):
    pass


class LibIdwarp:
    """The f2py module surface of libidwarp.pyf on the warp path."""

    def __init__(self, device=0, precision="fp64", nearWallFactor=1e-6, reusePlan=False):
        self._L = _lib.load()
        self._device = device
        self._extra = dict(precision=precision, nearWallFactor=nearWallFactor, reusePlan=reusePlan)
        # libidwarp.pyf:148-162 (gridinput), :164-169 (griddata), :204-206 (kd_tree)
        self.gridinput = _Module(aexp=3.0, bexp=5.0, alpha=0.25, ldeffact=1.0, symmtol=1e-6, errtol=5e-4,
                                 userotations=True, zerocornerrotations=True, cornerangle=30.0, evalmode=1,
                                 eval_exact=0, eval_fast=1)
        self.griddata = _Module(warpmeshdof=0, solvermeshdof=0, commonmeshdof=0)
        self.kd_tree = _Module(bucket_size=8)
        self._ctx = ctypes.c_void_p(0)
        self._npts = 0

    # -- context and options -------------------------------------------
    def _opts(self):
        o = _lib.IdwOptions()
        self._L.idw_default_options(ctypes.byref(o))
        g = self.gridinput
        o.aExp, o.bExp, o.LdefFact, o.alpha = g.aexp, g.bexp, g.ldeffact, g.alpha
        o.errTol, o.cornerAngle = g.errtol, g.cornerangle
        o.evalMode = _lib.EVAL_FAST if int(g.evalmode) == g.eval_fast else _lib.EVAL_EXACT
        o.useRotations, o.zeroCornerRotations = int(bool(g.userotations)), int(bool(g.zerocornerrotations))
        o.bucketSize = int(self.kd_tree.bucket_size)
        o.precision = _lib.PREC_MIXED if self._extra["precision"] == "mixed" else _lib.PREC_FP64
        o.nearWallFactor = float(self._extra["nearWallFactor"])
        o.reusePlan = int(bool(self._extra["reusePlan"]))
        return o

    def _ensure(self):
        if not self._ctx.value:
            _lib.check(self._L.idw_create(ctypes.byref(self._ctx), int(self._device), ctypes.byref(self._opts())))
        else:
            _lib.check(self._L.idw_set_options(self._ctx, ctypes.byref(self._opts())))
        return self._ctx

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            self._L.idw_destroy(ctx)
            self._ctx = None

    # -- set-up --------------------------------------------------------
    def initpetsc(self, comm=None):
        """libidwarp.pyf:10-12: nothing to initialise."""

    def createcommongrid(self, volnodes):
        """libidwarp.pyf:14-17 (createGrid.F90): volnodes is (3, n)."""
        v = np.ascontiguousarray(np.asarray(volnodes, dtype=np.float64).T)
        _lib.check(self._L.idw_set_volume_nodes(self._ensure(), _lib.ptr(v), v.size))
        self.griddata.warpmeshdof = self.griddata.commonmeshdof = v.size

    def pointreduce(self, pts, tol):
        """libidwarp.pyf:195-202: pts (3, n) -> uniquepts (3, n), link (n, 1-based), nunique."""
        p = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).T)
        uniq, link, nu = _lib.point_reduce(p, tol)
        out = np.zeros((3, len(p)))
        out[:, :nu] = uniq.T
        return out, link, nu

    def setsymmetryplanes(self, pts, normals):
        """libidwarp.pyf:69-73: pts and normals are (3, n)."""
        p = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(3, -1).T)
        n = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(3, -1).T)
        _lib.check(self._L.idw_set_symmetry_planes(self._ensure(), _lib.ptr(p), _lib.ptr(n), len(p)))

    def initializewarping(self, pts, uniquepts, link, facesizes, faceconn, restartfile=""):
        """libidwarp.pyf:55-68: pts flat 3*n, uniquepts (3, nunique), link
        1-based, faceconn 0-based into pts."""
        p = _lib.f64(pts)
        u = np.ascontiguousarray(np.asarray(uniquepts, dtype=np.float64).T)
        link, fs, fc = _lib.i32(link), _lib.i32(facesizes), _lib.i32(faceconn)
        restart = str(restartfile).strip() if restartfile is not None else ""
        _lib.check(self._L.idw_set_restart_file(self._ensure(), restart.encode() if restart else None))
        _lib.check(self._L.idw_initialize_warping(self._ensure(), _lib.ptr(p), p.size // 3, _lib.ptr(u), len(u),
                                                  _lib.ptr(link), _lib.ptr(fs), fs.size, _lib.ptr(fc), fc.size))
        self._npts = p.size // 3

    def deallocatepatchio(self):
        """libidwarp.pyf:25-26: no patch data is kept."""

    def setexternalmeshindices(self, solver_indices):
        """libidwarp.pyf:40-43: 0-based per-DOF indices."""
        ind = _lib.i32(solver_indices)
        _lib.check(self._L.idw_set_external_mesh_indices(self._ensure(), _lib.ptr(ind), ind.size))
        self.griddata.solvermeshdof = ind.size

    # -- the warp path -------------------------------------------------
    def setsurfacecoordinates(self, coordinates):
        c = _lib.f64(coordinates)
        _lib.check(self._L.idw_set_surface_coordinates(self._ensure(), _lib.ptr(c), c.size))

    def getsurfacecoordinates(self, coordinates):
        """intent(inout): fills ``coordinates`` (flat, 3*n)."""
        out = np.zeros(np.asarray(coordinates).size)
        _lib.check(self._L.idw_get_surface_coordinates(self._ensure(), _lib.ptr(out), out.size))
        coordinates[...] = out.reshape(np.shape(coordinates))

    def warpmesh(self):
        _lib.check(self._L.idw_warp_mesh(self._ensure()))

    def getvolumecoordinates(self, ndof):
        out = np.zeros(int(ndof))
        _lib.check(self._L.idw_get_volume_coordinates(self._ensure(), _lib.ptr(out), out.size))
        return out

    getcommonvolumecoordinates = getvolumecoordinates

    def warp_to_solver_grid(self, warp_grid, solver_grid):
        w = _lib.f64(warp_grid)
        s = np.zeros(np.asarray(solver_grid).size)
        _lib.check(self._L.idw_warp_to_solver_grid(self._ensure(), _lib.ptr(w), w.size, _lib.ptr(s), s.size))
        solver_grid[...] = s.reshape(np.shape(solver_grid))

    def solver_to_warp_grid(self, solver_grid, warp_grid):
        s = _lib.f64(solver_grid)
        w = np.zeros(np.asarray(warp_grid).size)
        _lib.check(self._L.idw_solver_to_warp_grid(self._ensure(), _lib.ptr(s), s.size, _lib.ptr(w), w.size))
        warp_grid[...] = w.reshape(np.shape(warp_grid))

    def warpderiv(self, dxv_f):
        d = _lib.f64(dxv_f)
        _lib.check(self._L.idw_warp_deriv(self._ensure(), _lib.ptr(d), d.size))

    def warpderivfwd(self, xsdot, outvec):
        """libidwarp.pyf:107-112: outvec (intent inout, 3*n_vol) = (dXv/dXs) xsdot."""
        d = _lib.f64(xsdot)
        out = np.zeros(np.asarray(outvec).size)
        _lib.check(self._L.idw_warp_deriv_fwd(self._ensure(), _lib.ptr(d), d.size, _lib.ptr(out), out.size))
        outvec[...] = out.reshape(np.shape(outvec))

    def getdxs(self, output):
        """intent(inout): fills ``output`` (flat, 3*n)."""
        out = np.zeros(np.asarray(output).size)
        _lib.check(self._L.idw_get_dxs(self._ensure(), _lib.ptr(out), out.size))
        output[...] = out.reshape(np.shape(output))
