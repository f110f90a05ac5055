"""USMesh: IDWarp's unstructured-mesh API on the MI355X volume-warp library.

Method and option names, argument meaning and error behaviour follow
idwarp/UnstructuredMesh.py (class USMesh, :43).  The volume grid comes from
an OpenFOAM ASCII case, a multiblock PLOT3D file, or memory (``gridFile``
may be an (N, 3) array or a ``.npy`` path); CGNS files are not read.  Two
options are added: ``precision`` ("fp64" or "mixed") and ``nearWallFactor``.
All numerics run in HIP kernels through ``libidwarp_hip.so``; nothing here
computes a warp on the CPU.
"""
import ctypes

import numpy as np

from . import _lib


class Error(Exception):
    """Raised for user errors, as baseclasses.Error in the reference."""

    def __init__(self, message):
        super().__init__(message)
        self.message = message


_DEFAULTS = {
    "gridFile": [(str, np.ndarray, type(None)), None],
    "fileType": [str, ["memory", "NPY", "CGNS", "OpenFOAM", "PLOT3D"]],
    "specifiedSurfaces": [(list, type(None)), None],
    "symmetrySurfaces": [(list, type(None)), None],
    "symmetryPlanes": [(list, type(None)), None],
    "aExp": [float, 3.0],
    "bExp": [float, 5.0],
    "LdefFact": [float, 1.0],
    "alpha": [float, 0.25],
    "errTol": [float, 0.0005],
    "evalMode": [str, ["fast", "exact"]],
    "symmTol": [float, 1e-6],
    "useRotations": [bool, True],
    "zeroCornerRotations": [bool, True],
    "cornerAngle": [float, 30.0],
    "restartFile": [(str, type(None)), None],
    "bucketSize": [int, 8],
    # device-only
    "precision": [str, ["fp64", "mixed"]],
    "nearWallFactor": [float, 1e-6],
    "reusePlan": [bool, False],
    "autotune": [bool, False],
    "deterministic": [bool, False],
    "storeWeights": [bool, False],
    "device": [int, 0],
}


class USMesh:
    """Volume mesh warping (UnstructuredMesh.py:43-1091)."""

    def __init__(self, options=None, comm=None, debug=False):
        if options is None:
            raise Error("USMesh needs an options dictionary (options=...) when it is created")
        self.comm = comm
        self.name = "IDWarp"
        self.category = "Volume mesh warping"
        self._opts = {}
        for k, (typ, default) in _DEFAULTS.items():
            self._opts[k.lower()] = (k, default[0] if isinstance(default, list) else default)
        for k, v in options.items():
            self.setOption(k, v)
        L = _lib.load()
        self._L = L
        self._ctx = ctypes.c_void_p(0)
        _lib.check(L.idw_create(ctypes.byref(self._ctx), int(self.getOption("device")), ctypes.byref(self._copts())))
        self.dtype = "d"
        self.warpInitialized = False
        self.nSurf = 0
        self.nVol = 0
        self.OFData = {}
        self.plot3dSizes = None
        self.fileType = self.getOption("fileType")
        grid = self.getOption("gridFile")
        if grid is not None:
            if self.fileType == "OpenFOAM":
                # UnstructuredMesh.py:1004-1049 (_readOFGrid)
                from . import openfoam

                self.OFData = openfoam.read_polymesh(grid)
                grid = self.OFData["x0"]
            elif self.fileType == "PLOT3D":
                # UnstructuredMesh.py:121-122, readPlot3d.F90:1-162
                from . import plot3d

                self.plot3dSizes, grid = plot3d.read_grid(grid)
            elif isinstance(grid, str):
                if self.fileType not in ("NPY", "memory") or not grid.endswith(".npy"):
                    raise Error("Grid files other than OpenFOAM cases, PLOT3D files and .npy arrays are not "
                                "supported (the CGNS reader is not part of this library)")
                grid = np.load(grid)
            self.setVolumeNodes(grid)

    # ------------------------------------------------------------------
    # options (BaseSolver subset)
    # ------------------------------------------------------------------
    def setOption(self, name, value):
        key = name.lower()
        if key not in self._opts:
            raise Error("Option '%s' is not a valid IDWarp option" % name)
        spec = _DEFAULTS[self._opts[key][0]]
        if isinstance(spec[1], list) and value not in spec[1]:
            raise Error("Option '%s' must be one of %s" % (name, spec[1]))
        if spec[0] is float and isinstance(value, (int, np.integer)) and not isinstance(value, bool):
            value = float(value)
        old = self._opts[key]
        self._opts[key] = (old[0], value)
        if key in ("precision", "nearwallfactor", "aexp", "bexp", "alpha") and self.getOption("precision") == "mixed":
            floor = _lib.load().idw_mixed_near_wall_floor(self.getOption("aExp"), self.getOption("bExp"),
                                                          self.getOption("alpha"))
            if not self.getOption("nearWallFactor") >= floor:
                self._opts[key] = old
                raise Error("nearWallFactor %g is below the mixed-precision floor %.3g for aExp=%g, bExp=%g, "
                            "alpha=%g: FP32 weights of nearer pairs would overflow" %
                            (self.getOption("nearWallFactor"), floor, self.getOption("aExp"),
                             self.getOption("bExp"), self.getOption("alpha")))

    def getOption(self, name):
        key = name.lower()
        if key not in self._opts:
            raise Error("Option '%s' is not a valid IDWarp option" % name)
        return self._opts[key][1]

    def printOptions(self):
        for key in sorted(self._opts):
            print("%-22s: %s" % (self._opts[key][0], self._opts[key][1]))

    def _copts(self):
        o = _lib.IdwOptions()
        self._L.idw_default_options(ctypes.byref(o)) if hasattr(self, "_L") else _lib.load().idw_default_options(
            ctypes.byref(o))
        g = self.getOption
        o.aExp, o.bExp, o.LdefFact, o.alpha = g("aExp"), g("bExp"), g("LdefFact"), g("alpha")
        o.errTol, o.cornerAngle = g("errTol"), g("cornerAngle")
        o.evalMode = _lib.EVAL_FAST if g("evalMode") == "fast" else _lib.EVAL_EXACT
        o.useRotations, o.zeroCornerRotations = int(bool(g("useRotations"))), int(bool(g("zeroCornerRotations")))
        o.bucketSize = int(g("bucketSize"))
        o.precision = _lib.PREC_MIXED if g("precision") == "mixed" else _lib.PREC_FP64
        o.nearWallFactor = g("nearWallFactor")
        o.reusePlan = int(bool(g("reusePlan")))
        o.deterministic = int(bool(g("deterministic")))
        o.storeWeights = int(bool(g("storeWeights")))
        return o

    def _setMeshOptions(self):
        """UnstructuredMesh.py:1051-1069: push options before each warp."""
        _lib.check(self._L.idw_set_options(self._ctx, ctypes.byref(self._copts())))

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            self._L.idw_destroy(ctx)
            self._ctx = None

    # ------------------------------------------------------------------
    # grid and surface definition
    # ------------------------------------------------------------------
    def setVolumeNodes(self, xv0):
        """Set the undeformed volume grid (the CGNS reader's createCommonGrid,
        src/IO/createGrid.F90), an (N, 3) array."""
        xv0 = _lib.f64(xv0).reshape(-1, 3)
        _lib.check(self._L.idw_set_volume_nodes(self._ctx, _lib.ptr(xv0), xv0.size))
        self.nVol = len(xv0)

    def setSurfaceDefinition(self, pts, conn, faceSizes, cgnsBlockID=None):
        """UnstructuredMesh.py:196-262: point-reduce the surface, sort the
        unique points by (z, y, x), relink and initialise the warp."""
        pts = _lib.f64(pts).reshape(-1, 3)
        conn = _lib.i32(np.array(conn).flatten())
        faceSizes = _lib.i32(faceSizes)
        uniquePts, link, nUnique = _lib.point_reduce(pts, 1e-12)
        ind = np.lexsort(uniquePts.T)
        uniquePts = np.ascontiguousarray(uniquePts[ind])
        inv = np.zeros_like(ind)
        inv[ind] = np.arange(len(ind))
        link = _lib.i32(inv[link - 1] + 1)
        self._setSymmetryConditions()
        self._setMeshOptions()
        restart = self.getOption("restartFile")
        _lib.check(self._L.idw_set_restart_file(self._ctx, restart.encode() if restart else None))
        _lib.check(self._L.idw_initialize_warping(self._ctx, _lib.ptr(pts), len(pts), _lib.ptr(uniquePts),
                                                  len(uniquePts), _lib.ptr(link), _lib.ptr(faceSizes),
                                                  len(faceSizes), _lib.ptr(conn), len(conn)))
        self.nSurf = len(pts)
        self.warpInitialized = True

    def setSurfaceDefinitionFromFile(self, surfFile):
        """Surface definition from a multiblock PLOT3D surface file: every
        cell of every zone is a quadrilateral (UnstructuredMesh.py:263-294,
        readPlot3dSurface, readPlot3d.F90:164-262)."""
        from . import plot3d

        pts, conn, _ = plot3d.read_surface(surfFile)
        self.setSurfaceDefinition(pts, conn, 4 * np.ones(len(conn), dtype=np.int32))

    def setSurfaceFromFile(self, surfFile):
        """New surface coordinates from a PLOT3D surface file with the zones
        and sizes of the one given to setSurfaceDefinitionFromFile
        (UnstructuredMesh.py:296-323)."""
        from . import plot3d

        self.setSurfaceCoordinates(plot3d.read_surface(surfFile)[0])

    def _setSymmetryConditions(self):
        """UnstructuredMesh.py:752-962 for user-given planes: the
        'symmetryPlanes' option [[point, normal], ...].  Planes found from
        CGNS symmetry families ('symmetrySurfaces') need a CGNS reader, which
        this backend does not have."""
        planes = self.getOption("symmetryPlanes")
        if planes is None:
            if self.getOption("symmetrySurfaces") is not None:
                raise Error("Automatic determination of symmetry surfaces needs a CGNS grid. Please specify the "
                            "symmetry planes using the 'symmetryPlanes' option.")
            planes = []
        pts = np.ascontiguousarray([p for p, _ in planes], np.float64).reshape(-1, 3)
        normals = np.ascontiguousarray([n for _, n in planes], np.float64).reshape(-1, 3)
        _lib.check(self._L.idw_set_symmetry_planes(self._ctx, _lib.ptr(pts), _lib.ptr(normals), len(pts)))

    def setExternalMeshIndices(self, ind):
        """0-based per-DOF indices of the solver grid into the common grid."""
        ind = _lib.i32(ind)
        _lib.check(self._L.idw_set_external_mesh_indices(self._ctx, _lib.ptr(ind), ind.size))

    # ------------------------------------------------------------------
    # surface coordinates
    # ------------------------------------------------------------------
    def _setInternalSurface(self):
        """UnstructuredMesh.py:602-750: without setSurfaceDefinition(), an
        OpenFOAM case is warped about all of its patch, wall and slip
        boundary faces."""
        if self.warpInitialized:
            return
        if self.fileType != "OpenFOAM" or not self.OFData:
            raise Error("No surface defined: call setSurfaceDefinition() first")
        import warnings

        from . import openfoam

        warnings.warn("No surface was defined: warping about the grid's own patch, wall and slip faces. A solver "
                      "that drives this mesh should call setExternalMeshIndices() and setSurfaceDefinition() "
                      "before the first warp.", stacklevel=2)
        pts, conn, faceSizes = openfoam.warp_surface(self.OFData)
        self.setSurfaceDefinition(pts, conn, faceSizes)

    def getSurfaceCoordinates(self):
        self._setInternalSurface()
        coords = np.zeros((self.nSurf, 3), self.dtype)
        _lib.check(self._L.idw_get_surface_coordinates(self._ctx, _lib.ptr(coords), coords.size))
        return coords

    def setSurfaceCoordinates(self, coordinates):
        if len(coordinates) != self.nSurf:
            raise Error("setSurfaceCoordinates got %d surface points; the surface definition has %d" %
                        (len(coordinates), self.nSurf))
        c = _lib.f64(coordinates).ravel()
        _lib.check(self._L.idw_set_surface_coordinates(self._ctx, _lib.ptr(c), c.size))

    # ------------------------------------------------------------------
    # warp, grids, derivatives
    # ------------------------------------------------------------------
    def warpMesh(self):
        """Warp the volume to the surface set by setSurfaceCoordinates()."""
        self._setInternalSurface()
        self._setMeshOptions()
        # a search picks parameters by timing, and the sums' order follows
        # them: a deterministic mesh keeps its parameters unless tune() is called
        if self.getOption("autotune") and not self.getOption("deterministic"):
            _lib.check(self._L.idw_tune(self._ctx, 0))
        _lib.check(self._L.idw_warp_mesh(self._ctx))

    def tune(self, force=False):
        """Pick the volume kernels' tile parameters for this mesh at the
        current options by timing the candidates (idw_tune); the option
        autotune does this before the first warp at new options."""
        self._setInternalSurface()
        self._setMeshOptions()
        _lib.check(self._L.idw_tune(self._ctx, int(bool(force))))

    TUNING_NAMES = ("forwardChunk", "mergeWindow", "walkBuffer", "exactGroup", "exactChunkShift")
    TUNING_SLOTS = (("walkBuffer", (64, 128, 256)), ("forwardChunk", (16, 32, 64)), ("mergeWindow", (8, 16, 32)),
                    ("exactGroup", (16, 32, 64)), ("exactChunkShift", (-1, 0, 1)))

    def setTuning(self, **params):
        """Set tile parameters by name (forwardChunk, mergeWindow, walkBuffer,
        exactGroup, exactChunkShift) without a search."""
        chosen = self.getTuning()[0]
        unknown = set(params) - set(self.TUNING_NAMES)
        if unknown:
            raise Error("unknown tuning parameters: %s" % ", ".join(sorted(unknown)))
        chosen.update(params)
        p = (ctypes.c_int32 * 5)(*[int(chosen[k]) for k in self.TUNING_NAMES])
        _lib.check(self._L.idw_set_tuning(self._ctx, p))

    def getTuning(self):
        """(chosen parameters, {parameter: {candidate: ms per launch}}, state)
        with state "untuned", "current" or "stale" (tuned at other options)."""
        p = (ctypes.c_int32 * 5)()
        ms = (ctypes.c_double * 15)()
        rc = self._L.idw_get_tuning(self._ctx, p, ms, 15)
        if rc < 0:
            _lib.check(rc)
        chosen = dict(zip(self.TUNING_NAMES, list(p)))
        table, k = {}, 0
        for name, cands in self.TUNING_SLOTS:
            table[name] = {c: ms[k + i] for i, c in enumerate(cands) if ms[k + i] >= 0}
            k += len(cands)
        return chosen, table, ("untuned", "current", "stale")[rc]

    def _warpMeshAsync(self):
        """warpMesh enqueued on this mesh's stream; complete after _wait()."""
        self._setInternalSurface()
        self._setMeshOptions()
        _lib.check(self._L.idw_warp_mesh_async(self._ctx))

    def _warpDerivAsync(self, dXv, solverVec=True):
        """warpDeriv enqueued on this mesh's stream; complete after _wait()."""
        self._setMeshOptions()
        dXvWarp = self._warpOrder(dXv, solverVec)
        _lib.check(self._L.idw_warp_deriv_async(self._ctx, _lib.ptr(dXvWarp), dXvWarp.size))

    def _wait(self):
        _lib.check(self._L.idw_wait(self._ctx))

    def getWarpGrid(self):
        xv = np.zeros(3 * self.nVol, self.dtype)
        _lib.check(self._L.idw_get_volume_coordinates(self._ctx, _lib.ptr(xv), xv.size))
        return xv

    def getCommonGrid(self):
        return self.getWarpGrid()

    def writeGrid(self, fileName=None):
        """Write the current grid in the input's format
        (UnstructuredMesh.py:480-511): a PLOT3D grid to fileName, an
        OpenFOAM case's constant/polyMesh/points in place."""
        if self.fileType in ("CGNS", "PLOT3D"):
            if fileName is None:
                raise Error("writeGrid needs a fileName for a CGNS or PLOT3D grid")
            if self.fileType == "CGNS":
                raise Error("CGNS output is not part of this library")
            from . import plot3d

            plot3d.write_grid(fileName, self.plot3dSizes, self.getCommonGrid().reshape(-1, 3))
        elif self.fileType == "OpenFOAM":
            from . import openfoam

            openfoam.write_points(self.getOption("gridFile"), self.getCommonGrid().reshape(-1, 3))
        else:
            raise Error("writeGrid needs a CGNS, PLOT3D or OpenFOAM grid; this mesh was given in memory")

    def _solverDOF(self):
        n = ctypes.c_int64(0)
        _lib.check(self._L.idw_get_solver_dof(self._ctx, ctypes.byref(n)))
        return n.value

    def getSolverGrid(self):
        """Warped grid in the order of setExternalMeshIndices()."""
        warpGrid = self.getWarpGrid()
        solverGrid = np.zeros(self._solverDOF(), self.dtype)
        _lib.check(self._L.idw_warp_to_solver_grid(self._ctx, _lib.ptr(warpGrid), warpGrid.size, _lib.ptr(solverGrid),
                                                   solverGrid.size))
        return solverGrid

    def _warpOrder(self, dXv, solverVec):
        dXv = _lib.f64(dXv).ravel()
        if not solverVec:
            return dXv
        dXvWarp = np.zeros(3 * self.nVol, self.dtype)
        _lib.check(self._L.idw_solver_to_warp_grid(self._ctx, _lib.ptr(dXv), dXv.size, _lib.ptr(dXvWarp),
                                                   dXvWarp.size))
        return dXvWarp

    def warpDeriv(self, dXv, solverVec=True):
        """dXs = (dXv/dXs)^T dXv; the result is read with getdXs()."""
        self._setMeshOptions()
        dXvWarp = self._warpOrder(dXv, solverVec)
        _lib.check(self._L.idw_warp_deriv(self._ctx, _lib.ptr(dXvWarp), dXvWarp.size))

    def warpDerivFwd(self, dXs, solverVec=True):
        """Forward-mode product dXv = (dXv/dXs) dXs (UnstructuredMesh.py:428-459,
        warpDerivFwd.F90).  dXs has the shape of getSurfaceCoordinates();
        the result is in solver ordering when solverVec, else warp ordering.
        As the reference does, the mesh is warped first."""
        dXs = _lib.f64(dXs).ravel()
        if dXs.size != 3 * self.nSurf:
            raise Error("Incorrect dimensions of dXs: expected %d values, received %d" % (3 * self.nSurf, dXs.size))
        dXvWarp = np.zeros(3 * self.nVol, self.dtype)
        self.warpMesh()
        _lib.check(self._L.idw_warp_deriv_fwd(self._ctx, _lib.ptr(dXs), dXs.size, _lib.ptr(dXvWarp), dXvWarp.size))
        if solverVec:
            dXv = np.zeros(self._solverDOF(), self.dtype)
            _lib.check(self._L.idw_warp_to_solver_grid(self._ctx, _lib.ptr(dXvWarp), dXvWarp.size, _lib.ptr(dXv),
                                                       dXv.size))
            return dXv
        return dXvWarp

    def getdXs(self):
        dXs = np.zeros((self.nSurf, 3), self.dtype)
        _lib.check(self._L.idw_get_dxs(self._ctx, _lib.ptr(dXs), dXs.size))
        return dXs

    def verifyWarpDeriv(self, dXv=None, solverVec=True, dofStart=0, dofEnd=10, h=1e-6, randomSeed=314):
        """Central-FD check of warpDeriv on surface DOFs [dofStart, dofEnd)
        (src/warp/verifyWarpDeriv.F90:43-148).  Prints and returns the
        relative errors (%)."""
        if dXv is None:
            np.random.seed(randomSeed)
            dXvWarp = np.random.random(3 * self.nVol)
        elif solverVec:
            dXv = _lib.f64(dXv).ravel()
            dXvWarp = np.zeros(3 * self.nVol)
            _lib.check(self._L.idw_solver_to_warp_grid(self._ctx, _lib.ptr(dXv), dXv.size, _lib.ptr(dXvWarp),
                                                       dXvWarp.size))
        else:
            dXvWarp = _lib.f64(dXv).ravel()
        xs0 = self.getSurfaceCoordinates().ravel()
        self.warpMesh()
        self.warpDeriv(dXvWarp, solverVec=False)
        ad = self.getdXs().ravel()
        errs = []
        for dof in range(dofStart, min(dofEnd, xs0.size)):
            xs = xs0.copy()
            xs[dof] += h
            self.setSurfaceCoordinates(xs.reshape(-1, 3))
            self.warpMesh()
            xp = self.getWarpGrid()
            xs[dof] -= 2 * h
            self.setSurfaceCoordinates(xs.reshape(-1, 3))
            self.warpMesh()
            xm = self.getWarpGrid()
            fd = np.dot(dXvWarp, (xp - xm) / (2 * h))
            err = (fd - ad[dof]) / (ad[dof] if ad[dof] != 0 else 1.0) * 100
            print("Dof: %4d, FD: %18.10e, AD: %18.10e, Err(%%): %12.5e" % (dof, fd, ad[dof], err))
            errs.append(err)
        self.setSurfaceCoordinates(xs0.reshape(-1, 3))
        self.warpMesh()
        return np.array(errs)

    # ------------------------------------------------------------------
    # extras of the device library
    # ------------------------------------------------------------------
    def getLdef0(self):
        v = ctypes.c_double(0.0)
        _lib.check(self._L.idw_get_ldef0(self._ctx, ctypes.byref(v)))
        return v.value

    def getDenominator0(self):
        d = np.zeros(self.nVol)
        _lib.check(self._L.idw_get_denominator0(self._ctx, _lib.ptr(d), d.size))
        return d

    def getRestartInfo(self):
        """('none' | 'loaded' | 'saved', load-balancing costs per volume node)
        of the restart file used by the last initialisation."""
        st = ctypes.c_int32(0)
        _lib.check(self._L.idw_get_restart_info(self._ctx, ctypes.byref(st), None, 0))
        costs = np.zeros(self.nVol if st.value else 0)
        if st.value:
            _lib.check(self._L.idw_get_restart_info(self._ctx, ctypes.byref(st), _lib.ptr(costs), costs.size))
        return ["none", "loaded", "saved"][st.value], costs

    def getTimings(self):
        """ms of the last warp/deriv phases: [nodal+cluster, tree walk,
        forward kernel, reverse set-up, reverse kernel, nodal reverse]."""
        t = np.zeros(6)
        _lib.check(self._L.idw_get_timings(self._ctx, _lib.ptr(t)))
        return t

    def getDeviceMemory(self):
        """Bytes of device memory this mesh holds (idw_device_bytes): its
        surface, volume and plan buffers and, with storeWeights, the stored
        weights."""
        b = ctypes.c_int64()
        _lib.check(self._L.idw_device_bytes(self._ctx, ctypes.byref(b)))
        return int(b.value)

    def getFastCounts(self):
        """Fast-mode work: per-node sums of visits, accepts, leaf pairs, error
        checks, then per-wave sums of visits, accepting visits, bucket
        visits, bucket point evaluations."""
        c = np.zeros(8)
        self._setMeshOptions()
        _lib.check(self._L.idw_fast_counts(self._ctx, _lib.ptr(c)))
        return c

    @property
    def ctx(self):
        return self._ctx
