"""Several independent warp instances over one multi-component (overset) grid.

``MultiUSMesh`` mirrors idwarp.MultiUSMesh (idwarp/MultiUnstructuredMesh.py,
class MultiUSMesh, :47): the grid's blocks are grouped into zones by name,
every zone named in ``optionsDict`` becomes one ``USMesh`` (a nearfield
component, warped about its own surface) and every other zone is background
that keeps its initial nodes.  The solver's volume and surface vectors are
split among the instances by the CGNS-order volume node intervals and block
intervals of the zones, exactly as the reference does.

The reference takes a CGNS file written by ``cgns_utils combine`` and needs
cgns_utils to explode it.  Here the grid is given as blocks in that same
order, either in memory, ``[(blockName, nodes), ...]``, or as a multiblock
PLOT3D file with ``blockNames``.  A block's zone is its name without the
``.<digits>`` suffix that ``cgns_utils combine`` appends, and the blocks of a
zone must be contiguous, as the reference's block intervals assume
(:111-140).  Each instance warps on the GPU through ``USMesh``; only index
bookkeeping happens here.
"""
import re

import numpy as np

from .usmesh import USMesh, Error

_SUFFIX = re.compile(r"^(.*)\.\d+$")


def zone_name(blockName):
    """The zone a block belongs to: its name without a trailing .<digits>."""
    m = _SUFFIX.match(blockName)
    return m.group(1) if m else blockName


def _block_sizes(nodes):
    """(il, jl, kl) of a block given as a (kl, jl, il, 3) array (i fastest),
    else (N, 1, 1)."""
    a = np.asarray(nodes)
    if a.ndim == 4:
        return (a.shape[2], a.shape[1], a.shape[0])
    return (a.size // 3, 1, 1)


class _BackgroundZone:
    """A zone that is not warped: its nodes in CGNS order.  Plays the role of
    the reference's dummy background instance (:197-270), which it creates only
    to read these nodes back in solver order."""

    def __init__(self, xv0):
        self.xv0 = np.ascontiguousarray(xv0, np.float64).ravel()
        self.ind = None

    def setExternalMeshIndices(self, ind):
        self.ind = np.asarray(ind, np.int64)

    def getSolverGrid(self):
        return self.xv0[self.ind]

    def getCommonGrid(self):
        return self.xv0.copy()


class MultiUSMesh:
    """Independent deformation of several overset component meshes
    (MultiUnstructuredMesh.py:47-892)."""

    def __init__(self, CGNSFile, optionsDict, comm=None, dtype="d", debug=False, blockNames=None):
        """CGNSFile: the grid's blocks in CGNS order, as a list of
        (blockName, nodes) with nodes an (N, 3) array or a (kl, jl, il, 3)
        array, or the path of a multiblock PLOT3D file whose blocks are named
        by ``blockNames``.  optionsDict: {zoneName: USMesh options}; zones
        without an entry are background and are not warped (:53-72)."""
        if dtype != "d":
            raise Error("MultiUSMesh supports dtype 'd' only (there is no complex-step build of this library)")
        self.comm = comm
        self.dtype = dtype
        self.CGNSFile = CGNSFile
        blocks = self._read_blocks(CGNSFile, blockNames)

        # zones in order of appearance, each a contiguous run of blocks (:111-146)
        zoneNames, zoneBlocks = [], []
        for name, nodes in blocks:
            z = zone_name(name)
            if zoneNames and zoneNames[-1] == z:
                zoneBlocks[-1].append(nodes)
            elif z in zoneNames:
                raise Error("The blocks of zone '%s' are not contiguous; MultiUSMesh needs each zone's blocks in one "
                            "run, as cgns_utils combine writes them" % z)
            else:
                zoneNames.append(z)
                zoneBlocks.append([nodes])
        unknown = sorted(set(optionsDict) - set(zoneNames))
        if unknown:
            raise Error("optionsDict names zones that are not in the grid: %s" % ", ".join(unknown))

        self.zoneNames = zoneNames
        self.cgnsBlockIntervals, self.cgnsVolNodeIntervals, self.blockSizes = [], [], []
        blockCounter = nodeCounter = 0
        self.meshes, self.backgroundInstanceIDs = [], []
        for zoneNumber, (z, blks) in enumerate(zip(zoneNames, zoneBlocks)):
            xv0 = np.concatenate([np.asarray(b, np.float64).reshape(-1, 3) for b in blks])
            self.cgnsBlockIntervals.append([blockCounter, blockCounter + len(blks)])
            self.cgnsVolNodeIntervals.append([nodeCounter, nodeCounter + xv0.size])
            self.blockSizes.append([_block_sizes(b) for b in blks])
            blockCounter += len(blks)
            nodeCounter += xv0.size
            if z in optionsDict:
                opts = dict(optionsDict[z])
                opts["gridFile"] = xv0
                opts["fileType"] = "memory"
                self.meshes.append(USMesh(options=opts, comm=comm, debug=debug))
            else:
                self.meshes.append(_BackgroundZone(xv0))
                self.backgroundInstanceIDs.append(zoneNumber)

        self.numSurfNodes = None
        self.numVolNodes = None
        self.cgnsVolNodeMasks = []
        self.filtered2fullMaps = None
        self.defaultSolverGrid = None

    @staticmethod
    def _read_blocks(src, blockNames):
        if isinstance(src, str):
            if src.lower().endswith(".cgns"):
                raise Error("CGNS files are not read by this library; give the blocks in memory or as a multiblock "
                            "PLOT3D file with blockNames")
            from . import plot3d

            sizes, nodes = plot3d.read_grid(src)
            if blockNames is None or len(blockNames) != len(sizes):
                raise Error("A PLOT3D grid needs blockNames, one per block (%d blocks)" % len(sizes))
            out, off = [], 0
            for name, (il, jl, kl) in zip(blockNames, sizes):
                n = int(il * jl * kl)
                out.append((name, nodes[off:off + n].reshape(int(kl), int(jl), int(il), 3)))
                off += n
            return out
        return [(str(name), nodes) for name, nodes in src]

    # ------------------------------------------------------------------
    def _check_surface(self):
        if self.numSurfNodes is None:
            raise NameError("MultiUSMesh has no surface yet: call setSurfaceDefinition() first")

    def getSurfaceCoordinates(self):
        """All surface coordinates in solver ordering (:287-312); points of
        no instance are zero."""
        self._check_surface()
        pts = np.zeros((self.numSurfNodes, 3), dtype=self.dtype)
        for instanceID, mesh in enumerate(self.meshes):
            pts[self.filtered2fullMaps[instanceID], :] = mesh.getSurfaceCoordinates()
        return pts

    def setSurfaceCoordinates(self, pts):
        """Surface coordinates in solver ordering, shaped like
        getSurfaceCoordinates() (:314-334)."""
        self._check_surface()
        pts = np.asarray(pts, dtype=self.dtype).reshape(-1, 3)
        for instanceID, mesh in enumerate(self.meshes):
            mesh.setSurfaceCoordinates(pts[self.filtered2fullMaps[instanceID], :])

    def setExternalMeshIndices(self, ind):
        """0-based indices of the solver's volume DOFs into the CGNS-order
        grid (:336-414).  Each instance receives the indices that fall in its
        zone's interval, shifted to its own numbering; background zones fill
        the default solver grid and are dropped."""
        ind = np.asarray(ind, dtype=np.int64).ravel()
        if self.backgroundInstanceIDs is None:
            raise Error("setExternalMeshIndices can be called once per MultiUSMesh")
        self.numVolNodes = len(ind)
        self.defaultSolverGrid = np.zeros(self.numVolNodes, dtype=self.dtype)
        self.cgnsVolNodeMasks = []
        for instanceID, mesh in enumerate(self.meshes):
            start, end = self.cgnsVolNodeIntervals[instanceID]
            mask = (ind >= start) & (ind < end)
            self.cgnsVolNodeMasks.append(mask)
            mesh.setExternalMeshIndices(ind[mask] - start)
            if instanceID in self.backgroundInstanceIDs:
                self.defaultSolverGrid[mask] = mesh.getSolverGrid()
        for instanceID in reversed(range(len(self.meshes))):
            if instanceID in self.backgroundInstanceIDs:
                del self.meshes[instanceID]
                del self.cgnsVolNodeMasks[instanceID]
                del self.cgnsVolNodeIntervals[instanceID]
                del self.cgnsBlockIntervals[instanceID]
                del self.blockSizes[instanceID]
                del self.zoneNames[instanceID]
        self.backgroundInstanceIDs = None

    def getSolverGrid(self):
        """The grid in solver order (:416-443): background nodes as given,
        every instance's nodes as warped."""
        if self.defaultSolverGrid is None:
            raise Error("Run setExternalMeshIndices before getSolverGrid")
        solverGrid = self.defaultSolverGrid.copy()
        for instanceID, mesh in enumerate(self.meshes):
            solverGrid[self.cgnsVolNodeMasks[instanceID]] = mesh.getSolverGrid()
        return solverGrid

    def setSurfaceDefinition(self, pts, conn=None, faceSizes=None, cgnsBlockIDs=None, distTol=1e-6):
        """Split the solver surface among the instances by the 0-based CGNS
        block of each element (:445-594).  An instance receives its elements
        and the points they use, renumbered in increasing solver order;
        elements of background blocks go to no instance.  distTol is unused,
        as in the reference."""
        if self.backgroundInstanceIDs is not None:
            raise NameError("MultiUSMesh.setSurfaceDefinition() needs setExternalMeshIndices() to have run first")
        pts = np.asarray(pts, dtype=self.dtype).reshape(-1, 3)
        faceSizes = np.asarray(faceSizes, dtype=np.int64).ravel()
        conn = np.asarray(conn, dtype=np.int64).ravel()
        cgnsBlockIDs = np.asarray(cgnsBlockIDs).ravel()
        numElems = len(faceSizes)
        if len(cgnsBlockIDs) != numElems:
            raise ValueError("cgnsBlockIDs holds %d entries for %d elements" % (len(cgnsBlockIDs), numElems))
        if faceSizes.sum() != len(conn):
            raise ValueError("conn holds %d entries, faceSizes sum to %d" % (len(conn), int(faceSizes.sum())))
        self.numSurfNodes = pts.shape[0]
        self.instanceIDs = np.full(numElems, -1.0)
        for ii in range(len(self.meshes)):
            lo, hi = self.cgnsBlockIntervals[ii]
            self.instanceIDs[(cgnsBlockIDs >= lo) & (cgnsBlockIDs < hi)] = ii
        connInstance = np.repeat(self.instanceIDs, faceSizes)
        self.filtered2fullMaps = []
        for ii, mesh in enumerate(self.meshes):
            currConn = conn[connInstance == ii]
            currFaceSizes = faceSizes[self.instanceIDs == ii].astype(np.intc)
            filtered2fullMap = np.unique(currConn)
            currPts = pts[filtered2fullMap, :]
            currConn = np.searchsorted(filtered2fullMap, currConn).astype(np.intc)
            mesh.setSurfaceDefinition(currPts, currConn, currFaceSizes)
            self.filtered2fullMaps.append(filtered2fullMap)

    def getWarpGrid(self):
        """(volNodesList, numCoorTotal): each instance's grid as a flat array
        and the total coordinate count (:596-636)."""
        volNodesList = [mesh.getCommonGrid() for mesh in self.meshes]
        return volNodesList, int(sum(len(v) for v in volNodesList))

    def getdXs(self):
        """Surface seeds of the last warpDeriv summed over the instances
        (:638-665)."""
        self._check_surface()
        dXs = np.zeros((self.numSurfNodes, 3), dtype=self.dtype)
        for instanceID, mesh in enumerate(self.meshes):
            m = self.filtered2fullMaps[instanceID]
            dXs[m] = dXs[m] + mesh.getdXs()
        return dXs

    def warpMesh(self):
        """Warp every instance (:667-699).  The reference warps them one after
        another; here every instance's warp is enqueued on its own stream
        first and then all are waited for, so they run side by side on the
        device."""
        for mesh in self.meshes:
            mesh._warpMeshAsync()
        for mesh in self.meshes:
            mesh._wait()

    def warpDeriv(self, dXv, solverVec=True):
        """dXs = (dXv/dXs)^T dXv over all instances (:701-755): each instance
        takes the entries of dXv in its zone; the instances' reverse products
        run side by side, as in warpMesh."""
        dXv = np.asarray(dXv, dtype=self.dtype).ravel()
        for instanceID, mesh in enumerate(self.meshes):
            mesh._warpDerivAsync(dXv[self.cgnsVolNodeMasks[instanceID]], solverVec=solverVec)
        for mesh in self.meshes:
            mesh._wait()
        return self.getdXs()

    def warpDerivFwd(self, dXs):
        """dXv = (dXv/dXs) dXs in solver ordering (:757-814); background
        entries are zero."""
        self._check_surface()
        dXs = np.asarray(dXs, dtype=self.dtype).reshape(-1, 3)
        dXv = np.zeros(self.numVolNodes, dtype=self.dtype)
        for instanceID, mesh in enumerate(self.meshes):
            dXv[self.cgnsVolNodeMasks[instanceID]] = mesh.warpDerivFwd(dXs[self.filtered2fullMaps[instanceID]])
        return dXv

    def writeGrid(self, baseName=None):
        """Write each instance's grid (:819-861) as a multiblock PLOT3D file
        <baseName>_inst%03d.xyz with its zone's blocks (the reference writes
        CGNS, which this library does not)."""
        from . import plot3d

        if baseName is None:
            baseName = "IDWarpMulti"
        names = []
        for instanceID, mesh in enumerate(self.meshes):
            fileName = baseName + "_inst%03d.xyz" % instanceID
            plot3d.write_grid(fileName, self.blockSizes[instanceID], mesh.getCommonGrid().reshape(-1, 3))
            names.append(fileName)
        return names
