"""Python API of PASIf, kept call-for-call compatible with the reference's
``PASIf.py`` (``coo_tensor`` at ``PASIf.py:32-146``, ``PASIf`` at
``PASIf.py:148-633``) so that user scripts run unchanged, on top of the
MI355X driver in :mod:`pasif_amd.PASIfgpu`.

Behaviour kept from the reference:

* inputs are lists of per-system ``dia_matrix`` M, ``coo_matrix`` B and K,
  ``coo_tensor`` Gamma / Lambda (/ Psi for the adjoint), dense force pattern
  and initial conditions; the same assertions guard them (``PASIf.py:213-221,
  292-305, 439-472``);
* the systems are assembled block-diagonally (``475-543``) and scaled by
  ``-M^{-1}`` (``607-633``); ``A = (B + K)`` goes to the driver as sorted CSR
  (``250-255``);
* ``getTrajectory`` returns the ``(N*P + 1) x floor(T/saveSteps)`` array with
  the time row ``t*saveSteps/sampleRate`` on top (``404-421``).

Differences (documented, deliberate):

* assembly is vectorised (O(nnz), not the reference's O(S^2) loop) and
  accepts systems of different sizes;
* user inputs are never mutated (the reference rebinds
  ``vecForcePattern[0]`` and concatenates in place into ``vecGamma[0]``,
  ``PASIf.py:520-521, 584-586``);
* getters return numpy arrays; ``getGradient(save=0)`` accepts the argument
  the reference's examples pass;
* the GPU is chosen from ``LOCAL_RANK`` (default 0) instead of the mpi4py rank
  (``PASIf.py:27-30``): mpi4py is not required;
* extensions: :meth:`PASIf.setSystemExcitations` (one excitation and one step
  count per system instead of the Cartesian product), :meth:`setPrecision`,
  :meth:`setSolver`, :meth:`setTuning`, :meth:`getMaxAmplitudes`.
"""
from __future__ import annotations

import copy as cp
import os
from typing import List, Optional, Sequence

import numpy as np
from scipy.sparse import coo_matrix, dia_matrix, issparse

from .PASIfgpu import __GpuDriver as _GpuDriver

__all__ = ["coo_tensor", "PASIf"]


def _gpu_id() -> int:
    for k in ("LOCAL_RANK", "OMPI_COMM_WORLD_LOCAL_RANK", "SLURM_LOCALID"):
        if k in os.environ:
            try:
                return int(os.environ[k])
            except ValueError:
                pass
    return 0


class coo_tensor:  # noqa: N801  (reference name)
    """Sparse tensor in COO format with interleaved indices
    (``PASIf.py:32-146``): entry ``t`` has indices
    ``indices[t*ndim : (t+1)*ndim]``, the output (first) axis first."""

    def __init__(self, dimensions_: Optional[List[int]] = None):
        self.dimensions = list(dimensions_) if dimensions_ is not None else [2, 2]
        self.val = []
        self.indices = []

    def __str__(self) -> str:
        nd = len(self.dimensions)
        v = np.asarray(self.val, dtype=np.float64).reshape(-1)
        ix = np.asarray(self.indices, dtype=np.int64).reshape(-1, nd) if v.size else \
            np.zeros((0, nd), dtype=np.int64)
        lines = ["val: " + " ".join(str(x) for x in v)]
        for d in range(nd):
            lines.append(f"dim{d + 1}: " + " ".join(str(x) for x in ix[:, d]))
        return "\n" + "\n".join(lines) + "\n"

    # numpy views of the entries
    def _arrays(self):
        nd = len(self.dimensions)
        v = np.asarray(self.val, dtype=np.float64).reshape(-1)
        ix = (np.asarray(self.indices, dtype=np.int64).reshape(-1, nd) if v.size
              else np.zeros((0, nd), dtype=np.int64))
        return ix, v

    def concatenateTensor(self, tensor_: "coo_tensor"):
        """Append ``tensor_`` with every axis offset by this tensor's size
        along it (``PASIf.py:86-100``)."""
        if len(self.dimensions) != len(tensor_.dimensions):
            raise Exception("The number of dimensions of the two tensors are not the same")
        ix1, v1 = self._arrays()
        ix2, v2 = tensor_._arrays()
        ix2 = ix2 + np.asarray(self.dimensions, dtype=np.int64)[None, :]
        self.val = np.concatenate([v1, v2])
        self.indices = np.concatenate([ix1, ix2], axis=0).reshape(-1)
        self.dimensions = [a + b for a, b in zip(self.dimensions, tensor_.dimensions)]

    def multiplyByDiagMatrix(self, diagMatrix_):
        """``val[t] *= diag[first index of t]`` (``PASIf.py:102-115``)."""
        ix, v = self._arrays()
        if v.size:
            self.val = v * np.asarray(diagMatrix_, dtype=np.float64)[ix[:, 0]]

    def getIndicesDim(self, dim_: int):
        ix, _ = self._arrays()
        return list(ix[:, dim_])

    def isSquare(self) -> bool:
        return all(d == self.dimensions[0] for d in self.dimensions)

    def extendDimmensions(self, extension_: int):
        self.dimensions = [d + extension_ for d in self.dimensions]

    def offsetDimmension(self, targetDim_: int, offset_: int):
        ix, v = self._arrays()
        ix = ix.copy()
        ix[:, targetDim_] += offset_
        self.indices = ix.reshape(-1)


def _cat_tensors(tensors: Sequence[coo_tensor]) -> coo_tensor:
    """Block-diagonal concatenation of a list of tensors in one pass
    (``PASIf.py:504-519`` / ``584-586`` without mutating the inputs)."""
    nd = len(tensors[0].dimensions)
    for t in tensors:
        if len(t.dimensions) != nd:
            raise Exception("The number of dimensions of the two tensors are not the same")
    dims = np.array([t.dimensions for t in tensors], dtype=np.int64).reshape(len(tensors), nd)
    offs = np.concatenate([np.zeros((1, nd), dtype=np.int64), np.cumsum(dims, axis=0)[:-1]])
    # every tensor's entries gathered as they are, offsets added once
    idx, val = [], []
    for t in tensors:
        v = np.asarray(t.val, dtype=np.float64).reshape(-1)
        val.append(v)
        idx.append(np.asarray(t.indices).reshape(-1) if v.size else np.zeros(0, dtype=np.int64))
    out = coo_tensor(list(int(x) for x in dims.sum(axis=0)))
    out.val = np.concatenate(val) if val else np.zeros(0)
    counts = np.array([v.size for v in val], dtype=np.int64)
    ix = (np.concatenate(idx).astype(np.int64, copy=False).reshape(-1, nd) if idx
          else np.zeros((0, nd), dtype=np.int64))
    out.indices = (ix + np.repeat(offs, counts, axis=0)).reshape(-1)
    return out


def _coo_parts(m):
    """(row, col, value) arrays of one matrix, in the entry order
    ``coo_matrix(m)`` gives, read straight from COO / CSR / CSC storage
    (building a scipy object per system dominates the assembly of large
    batches)."""
    if issparse(m):
        fmt = m.format
        if fmt == "coo":
            return m.row, m.col, m.data
        if fmt == "csr":
            return np.repeat(np.arange(m.shape[0]), np.diff(m.indptr)), m.indices, m.data
        c = m.tocoo()
        return c.row, c.col, c.data
    a = np.asarray(m, dtype=np.float64)
    r, c = np.nonzero(a)
    return r, c, a[r, c]


def _diagonal(m) -> np.ndarray:
    """Main diagonal of one matrix (a DIA matrix's offset-0 row read
    directly)."""
    if issparse(m) and m.format == "dia":
        k = np.nonzero(np.asarray(m.offsets) == 0)[0]
        if k.size == 1:
            n = min(m.shape)
            d = np.zeros(n, dtype=np.float64)
            row = np.asarray(m.data[k[0]], dtype=np.float64)[:n]
            d[:row.size] = row
            return d
    return np.asarray(m.diagonal(), dtype=np.float64)


def _cat_matrices(mats, offs, n):
    rows, cols, vals = [], [], []
    for m in mats:
        r, c, v = _coo_parts(m)
        rows.append(r)
        cols.append(c)
        vals.append(v)
    # offsets added once for all blocks
    shift = np.repeat(np.asarray(offs, dtype=np.int64), [len(v) for v in vals])
    return coo_matrix((np.concatenate(vals).astype(np.float64, copy=False),
                       (np.concatenate(rows).astype(np.int64, copy=False) + shift,
                        np.concatenate(cols).astype(np.int64, copy=False) + shift)),
                      shape=(n, n))


def _neg_inverse_diag(M: dia_matrix) -> dia_matrix:
    """``-1 * inv(M)`` for a diagonal M (``PASIf.py:609, 627``)."""
    d = np.asarray(M.diagonal(), dtype=np.float64)
    if np.any(d == 0):
        raise ValueError("The matrix M must be invertible (zero on its diagonal).")
    n = d.size
    return dia_matrix((-(1.0 / d), [0]), shape=(n, n))


def assemble_arrays(vecM, vecB, vecK, vecGamma, vecLambda, vecForce, vecQ0, vecPsi=None,
                    adjoint=False):
    """Driver inputs of a list of systems, exactly as ``PASIf.setSystems`` /
    ``setJacobian`` hand them to the driver: block-diagonal unfold
    (``PASIf.py:475-605``), ``-M^{-1}`` scaling (``607-633``; the adjoint
    force pattern unscaled), sorted CSR of ``B + K`` (``250-255``).  Returns a
    dict with ``csr`` (scipy), ``gamma`` / ``lam`` / ``psi`` (coo_tensor),
    ``force``, ``q0`` and ``sizes``."""
    sizes = np.array([m.shape[0] for m in vecM], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    n = int(sizes.sum())
    diag = np.concatenate([_diagonal(m) for m in vecM])
    Dm = _neg_inverse_diag(dia_matrix((diag, [0]), shape=(n, n))).tocsc()
    Bs = coo_matrix(Dm.dot(_cat_matrices(vecB, offs, n)))
    Ks = coo_matrix(Dm.dot(_cat_matrices(vecK, offs, n)))
    D = Dm.diagonal()
    G = _cat_tensors(vecGamma)
    L = _cat_tensors(vecLambda)
    G.multiplyByDiagMatrix(D)
    L.multiplyByDiagMatrix(D)
    P = None
    if vecPsi is not None:
        P = _cat_tensors(vecPsi)
        P.multiplyByDiagMatrix(D)
    f = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in vecForce])
    if not adjoint:
        f = f * (-1.0 * D)
    q0 = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in vecQ0])
    csr = (Bs + Ks).tocsr()
    csr.sort_indices()
    return {"csr": csr, "gamma": G, "lam": L, "psi": P, "force": f, "q0": q0, "sizes": sizes,
            "B": Bs, "K": Ks}


class PASIf(_GpuDriver):
    """``PASIf(excitationSet, sampleRate, numsteps_=0, dCompute_=False,
    dSystem_=False, dSolver_=False)`` -- ``PASIf.py:148-194``."""

    def __init__(self, excitationSet, sampleRate: int, numsteps_: int = 0,
                 dCompute_: bool = False, dSystem_: bool = False, dSolver_: bool = False):
        if numsteps_ == 0:
            self.numsteps = len(excitationSet[0])
        else:
            self.numsteps = numsteps_
        super().__init__(excitationSet, sampleRate, self.numsteps, dCompute_, dSystem_, dSolver_,
                         _gpu_id())
        self.sampleRate = sampleRate
        self.globalSystemSize = 0
        self.globalAdjointSize = 0
        self.interpolSize = 0
        self.saveSteps = 1
        self.n_excitations = len(excitationSet)

        self.systemSet = False
        self.numberOfSystems = 0
        self.system_M = None
        self.system_B = None
        self.system_K = None
        self.system_Gamma = None
        self.system_Lambda = None
        self.system_forcePattern = None
        self.system_initialConditions = None
        self.system_sizes = None

        self.jacobianSet = False
        self.jacobian_M = None
        self.jacobian_B = None
        self.jacobian_K = None
        self.jacobian_Gamma = None
        self.jacobian_Lambda = None
        self.jacobian_forcePattern = None
        self.jacobian_initialConditions = None
        self.jacobian_Psi = None
        self.jacobian_sizes = None
        self._sys_exc = None
        self._sys_steps = None

    # ------------------------------------------------------------ setters
    def setExcitations(self, excitationSet, sampleRate: int):
        """``PASIf.py:196-200``"""
        self.sampleRate = sampleRate
        rc = self._loadExcitationsSet(excitationSet, self.sampleRate)
        if rc == 0:
            self.n_excitations = len(excitationSet)
        return rc

    def setSystems(self, vecM, vecB, vecK, vecGamma, vecLambda, vecForcePattern,
                   vecInitialConditions):
        """``PASIf.py:202-277``"""
        self.systemSet = False
        self.jacobianSet = False
        self.jacobian_sizes = None
        assert len(vecM) > 0
        assert len(vecM) == len(vecB) == len(vecK) == len(vecGamma) == len(vecLambda) == \
            len(vecForcePattern) == len(vecInitialConditions), \
            "The number of Systems in the input vectors must be the same."
        assert (type(vecM[0]) == dia_matrix and type(vecB[0]) == coo_matrix and
                type(vecK[0]) == coo_matrix and isinstance(vecGamma[0], coo_tensor) and
                isinstance(vecLambda[0], coo_tensor) and
                type(vecForcePattern[0]) == np.ndarray and
                type(vecInitialConditions[0]) == np.ndarray), "Inputs vectors types error."
        self.globalSystemSize = 0
        self.__checkInputs(vecM, vecB, vecK, vecGamma, vecLambda, vecForcePattern,
                           vecInitialConditions)
        self.__unfoldSystems(vecM, vecB, vecK, vecGamma, vecLambda, vecForcePattern,
                             vecInitialConditions)
        self.__systemPreprocessing()
        csr_K = (self.system_B + self.system_K).tocsr()
        csr_K.sort_indices()
        self._setFwdK(self.system_K.shape, csr_K.data, csr_K.indices, csr_K.indptr)
        self._setFwdGamma(self.system_Gamma.dimensions, self.system_Gamma.val,
                          self.system_Gamma.indices)
        self._setFwdLambda(self.system_Lambda.dimensions, self.system_Lambda.val,
                           self.system_Lambda.indices)
        self._setFwdPsi([1] * 5, [], [])
        self._setFwdForcePattern(self.system_forcePattern)
        self._setFwdInitialConditions(self.system_initialConditions)
        self._push_layout()
        self.systemSet = True
        self._allocateSystemOnDevice()

    def setJacobian(self, vecM, vecB, vecK, vecGamma, vecLambda, vecForcePattern,
                    vecInitialConditions, vecPsi):
        """``PASIf.py:279-369``"""
        self.jacobianSet = False
        self.globalAdjointSize = 0
        if not self.systemSet:
            raise ValueError("The system must be set before the jacobian")
        assert len(vecM) > 0
        assert len(vecM) == self.numberOfSystems, \
            "The number of jacobian must be the same as the number of systems."
        assert len(vecM) == len(vecB) == len(vecK) == len(vecGamma) == len(vecLambda) == \
            len(vecForcePattern) == len(vecInitialConditions), \
            "The number of Systems in the input vectors must be the same."
        assert (type(vecM[0]) == dia_matrix and type(vecB[0]) == coo_matrix and
                type(vecK[0]) == coo_matrix and isinstance(vecGamma[0], coo_tensor) and
                isinstance(vecLambda[0], coo_tensor) and
                type(vecForcePattern[0]) == np.ndarray and
                type(vecInitialConditions[0]) == np.ndarray and
                isinstance(vecPsi[0], coo_tensor)), "Inputs vectors types error."
        self.__checkInputs(vecM, vecB, vecK, vecGamma, vecLambda, vecForcePattern,
                           vecInitialConditions, vecPsi)
        self.__unfoldJacobians(vecM, vecB, vecK, vecGamma, vecLambda, vecForcePattern,
                               vecInitialConditions, vecPsi)
        self.__loadJacobian()

    def setJacobianBlocks(self, blocks):
        """``setJacobian`` for adjoint systems given already concatenated
        (``pasif_amd.adjoint.jacobian_blocks`` / ``energy_gradient_blocks``:
        the arrays ``setJacobian`` unfolds from per-system inputs, entry for
        entry), without a scipy object per system -- an extension for large
        batches."""
        self.jacobianSet = False
        self.globalAdjointSize = 0
        if not self.systemSet:
            raise ValueError("The system must be set before the jacobian")
        sizes = np.asarray(blocks.sizes, dtype=np.int64)
        assert sizes.size == self.numberOfSystems, \
            "The number of jacobian must be the same as the number of systems."
        assert np.array_equal(np.asarray(blocks.fwd_sizes), np.asarray(self.system_sizes)), \
            "The adjoint blocks were built for systems of other sizes."
        n = int(sizes.sum())
        self.globalAdjointSize = n
        self.jacobian_sizes = sizes
        self.jacobian_M = dia_matrix((-np.ones(n), [0]), shape=(n, n))
        self.jacobian_B = coo_matrix((n, n))
        self.jacobian_K = blocks.K
        cp = []
        for t in (blocks.gamma, blocks.lam, blocks.psi):
            c = coo_tensor(list(t.dimensions))
            c.val, c.indices = t.val, t.indices
            cp.append(c)
        self.jacobian_Gamma, self.jacobian_Lambda, self.jacobian_Psi = cp
        self.jacobian_forcePattern = np.zeros(n)
        self.jacobian_initialConditions = np.asarray(blocks.init, dtype=np.float64)
        self.__loadJacobian()

    def __loadJacobian(self):
        """Scale the unfolded adjoint system and hand it to the driver."""
        self.__jacobianPreprocessing()
        jac_csr_K = (self.jacobian_B + self.jacobian_K).tocsr()
        jac_csr_K.sort_indices()
        self._setBwdK(self.jacobian_K.shape, jac_csr_K.data, jac_csr_K.indices, jac_csr_K.indptr)
        self._setBwdGamma(self.jacobian_Gamma.dimensions, self.jacobian_Gamma.val,
                          self.jacobian_Gamma.indices)
        self._setBwdLambda(self.jacobian_Lambda.dimensions, self.jacobian_Lambda.val,
                           self.jacobian_Lambda.indices)
        self._setBwdForcePattern(self.jacobian_forcePattern)
        self._setBwdInitialConditions(self.jacobian_initialConditions)
        self._setBwdPsi(self.jacobian_Psi.dimensions, self.jacobian_Psi.val,
                        self.jacobian_Psi.indices)
        self._push_layout()
        self.jacobianSet = True
        self._allocateAdjointOnDevice()

    def setInterpolationMatrix(self, interpolationMatrix_):
        """``PASIf.py:371-385``; the kernels read rows 0 and 1, so at least two
        rows are required."""
        for i in range(len(interpolationMatrix_)):
            if len(interpolationMatrix_[i]) % 2 != 0:
                raise ValueError("The windows size must be even.")
            elif len(interpolationMatrix_[i]) != len(interpolationMatrix_[0]):
                raise ValueError("The windows size must be the same for all the rows.")
        if len(interpolationMatrix_) < 2:
            raise ValueError("The interpolation matrix needs at least two rows.")
        self.interpolSize = len(interpolationMatrix_)
        self.interpolationMatrix = np.array(interpolationMatrix_, dtype=np.float64).flatten()
        self._setInterpolationMatrix(self.interpolationMatrix, len(interpolationMatrix_[0]))

    def setModulationBuffer(self, modulationBuffer_):
        """``PASIf.py:387-396`` (stored; no kernel reads it, as in the
        reference)."""
        if len(modulationBuffer_) == 0:
            raise ValueError("The modulation buffer must be non-empty.")
        elif len(modulationBuffer_) > 32000:
            raise ValueError("The modulation buffer don't fit in the GPU cst memory.")
        self.modulationBuffer = modulationBuffer_
        self._setModulationBuffer(self.modulationBuffer)

    # ------------------------------------------------------------ getters
    def getAmplitudes(self):
        """Final state of every DOF, ``[replica][dof]`` (``PASIf.py:398-402``)."""
        if not self.systemSet:
            raise ValueError("The system must be set before computing the amplitudes.")
        return self._getAmplitudes()

    def getTrajectory(self, saveSteps: int = 1):
        """``PASIf.py:404-421``: row 0 is time, rows 1.. the DOFs."""
        if not self.systemSet:
            raise ValueError("The system must be set before computing the trajectory.")
        self.saveSteps = saveSteps
        trajectory = self._getTrajectory(saveSteps)
        n_out = self._out_len()
        n_saved = trajectory.size // n_out if n_out else 0
        out = np.zeros((n_out + 1, n_saved))
        out[0] = np.arange(n_saved) * self.saveSteps / self.sampleRate
        out[1:] = trajectory.reshape(n_saved, n_out).T
        return out

    def getGradient(self, save: int = 0):
        """lambda_0 of the adjoint, ``[replica][adjoint dof]``
        (``PASIf.py:423-429``)."""
        if not self.jacobianSet:
            raise ValueError("The jacobian must be set before computing the gradient.")
        return self._getGradient(save)

    def getMatrices(self):
        """``PASIf.py:431-434``"""
        if not self.systemSet:
            raise ValueError("System is not set yet")
        K_assembled = (self.system_B + self.system_K)
        return (K_assembled, self.system_Gamma, self.system_Lambda, self.system_forcePattern,
                self.system_initialConditions)

    # ------------------------------------------------------------ extensions
    def setSystemExcitations(self, excitation_of_system: Optional[Sequence[int]] = None,
                             steps_of_system: Optional[Sequence[int]] = None):
        """Drive system ``s`` by excitation ``excitation_of_system[s]`` only
        (-1: unforced) and integrate it ``steps_of_system[s]`` steps.  Outputs
        are then ``[dof]`` instead of ``[excitation][dof]``."""
        self._sys_exc = None if excitation_of_system is None else \
            np.asarray(excitation_of_system, dtype=np.int64)
        self._sys_steps = None if steps_of_system is None else \
            np.asarray(steps_of_system, dtype=np.int64)
        if self.systemSet:
            self._push_layout()

    def setPrecision(self, bits: int):
        self._setPrecision(bits)

    def setSolver(self, solver: str, steps_per_launch: int = 256, launches_per_graph: int = 16):
        """"persistent" (all steps in one launch), "plain" (one launch per
        step) or "graph" (launches of ``steps_per_launch`` steps replayed
        from a hipGraph of ``launches_per_graph`` launches)."""
        if solver == "graph":
            self._setGraphSlicing(steps_per_launch, launches_per_graph)
        self._setSolver(solver)

    def setTuning(self, tile_lanes: int = 0, slot_config: int = 0, checkpoint_every: int = 0):
        self._setTuning(tile_lanes, slot_config, checkpoint_every)

    def setKernel(self, kernel: str = "auto", lanes_per_system: int = 0, group_bias: int = 0):
        """"tile" (rows across a workgroup's lanes, systems of up to 1024
        lanes of row segments), "lane" (one lane -- or a group of lanes
        exchanging values by lane rotations -- per system, state in
        registers), "big" (one workgroup per system with every vector in
        device memory: any size, no trajectory or amplitude window) or
        "auto" (lane kernel for the structures that fit it, tiles for the
        rest, "big" for systems that fit neither).  ``lanes_per_system`` 0 lets a cost model pick
        the group size per structure (``group_bias`` moves every choice that
        many steps along the sizes that fit; the autotuner probes it); a
        power of two up to 64 requests that size."""
        self._setKernel(kernel)
        self._setLaneGroup(lanes_per_system)
        self._setLaneBias(group_bias)

    def setAdjointMode(self, mode: str = "reference"):
        """``"reference"``: the reference's adjoint recursion (continuous
        adjoint, ``src/__GpuDriver.cu:746-783``).  ``"exact"``: the exact
        discrete adjoint, the derivative of the computed RK4 map (tile and
        big kernels)."""
        self._setAdjointMode(mode)

    def kernelInfo(self, adjoint: bool = False) -> dict:
        """Kernel the loaded problem runs on, its structure classes and the
        checkpoint interval of the adjoint."""
        return self._kernelInfo(adjoint)

    def getMaxAmplitudes(self, window: int):
        """Final states and ``max |q_m|`` over the post-step states of the last
        ``window`` steps of every DOF, from one fused solve."""
        self._setAmplitudeWindow(window)
        try:
            q = self.getAmplitudes()
            return q, self._getMaxAmplitudes()
        finally:
            self._setAmplitudeWindow(0)

    # ------------------------------------------------------------ internals
    def _max_steps(self) -> int:
        if self._sys_steps is not None:
            return int(self._sys_steps.max())
        return int(self.numsteps)

    def _out_len(self) -> int:
        if self._sys_exc is not None:
            return self.globalSystemSize
        return self.globalSystemSize * self.n_excitations

    def _push_layout(self):
        self._setSystemsLayout(
            sizes_fwd=self.system_sizes,
            sizes_adj=self.jacobian_sizes,
            exc_of_system=self._sys_exc, steps_of_system=self._sys_steps)

    def __checkInputs(self, inputVecM, inputVecB, inputVecK, inputVecGamma, inputVecLambda,
                      inputVecForcePattern, inputVecInitialConditions, inputVecPsi=None):
        """``PASIf.py:439-472``"""
        for i in range(len(inputVecM)):
            ndofs = inputVecM[i].shape[0]
            assert inputVecM[i].shape[0] == inputVecM[i].shape[1], "The matrix M must be squared."
            assert inputVecB[i].shape[0] == inputVecB[i].shape[1], "The matrix B must be squared."
            assert inputVecK[i].shape[0] == inputVecK[i].shape[1], "The matrix K must be squared."
            if inputVecPsi is None:
                assert inputVecGamma[i].dimensions[0] == inputVecGamma[i].dimensions[1] == \
                    inputVecGamma[i].dimensions[2], "The System Gamma tensor must be squared."
                assert inputVecLambda[i].dimensions[0] == inputVecLambda[i].dimensions[1] == \
                    inputVecLambda[i].dimensions[2] == inputVecLambda[i].dimensions[3], \
                    "The System Lambda tensor must be squared."
            assert inputVecB[i].shape[0] == ndofs, "The matrix B must have the same size as the matrix M."
            assert inputVecK[i].shape[0] == ndofs, "The matrix K must have the same size as the matrix M."
            assert inputVecGamma[i].dimensions[0] == ndofs, "The Gamma tensor must have the same size as the matrix M."
            assert inputVecLambda[i].dimensions[0] == ndofs, "The Lambda tensor must have the same size as the matrix M."
            assert inputVecForcePattern[i].shape[0] == ndofs, "The force pattern vector must have the same size as the matrix M."
            assert inputVecInitialConditions[i].shape[0] == ndofs, "The initial conditions vector must have the same size as the matrix M."
            if inputVecPsi is not None:
                assert inputVecPsi[i].dimensions[0] == ndofs, "The Psi tensor must have the same size as the matrix M."
                assert inputVecPsi[i].dimensions[0] == inputVecPsi[i].dimensions[1], \
                    "The two highest dim of the Psi tensor should be of the size of the adjoint system."
                assert inputVecPsi[i].dimensions[2] == inputVecPsi[i].dimensions[3] == \
                    inputVecPsi[i].dimensions[4], \
                    "The tree smallest dim of the Psi tensor should be of the size of the forward system."

    def __unfoldSystems(self, inputVecM, inputVecB, inputVecK, inputVecGamma, inputVecLambda,
                        inputVecForcePattern, inputVecInitialConditions):
        """Block-diagonal assembly (``PASIf.py:475-543``), vectorised."""
        self.numberOfSystems = len(inputVecM)
        sizes = np.array([m.shape[0] for m in inputVecM], dtype=np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        n = int(sizes.sum())
        self.globalSystemSize = n
        self.system_sizes = sizes
        diag = np.concatenate([_diagonal(m) for m in inputVecM])
        self.system_M = dia_matrix((diag, [0]), shape=(n, n))
        self.system_B = _cat_matrices(inputVecB, offs, n)
        self.system_K = _cat_matrices(inputVecK, offs, n)
        self.system_Gamma = _cat_tensors(inputVecGamma)
        self.system_Lambda = _cat_tensors(inputVecLambda)
        self.system_forcePattern = np.concatenate(
            [np.asarray(f, dtype=np.float64).reshape(-1) for f in inputVecForcePattern])
        self.system_initialConditions = np.concatenate(
            [np.asarray(q, dtype=np.float64).reshape(-1) for q in inputVecInitialConditions])

    def __unfoldJacobians(self, inputVecM, inputVecB, inputVecK, inputVecGamma, inputVecLambda,
                          inputVecForcePattern, inputVecInitialConditions, inputVecPsi):
        """``PASIf.py:545-605``, vectorised and without mutating the inputs."""
        sizes = np.array([m.shape[0] for m in inputVecM], dtype=np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        n = int(sizes.sum())
        self.globalAdjointSize = n
        self.jacobian_sizes = sizes
        diag = np.concatenate([_diagonal(m) for m in inputVecM])
        self.jacobian_M = dia_matrix((diag, [0]), shape=(n, n))
        self.jacobian_B = _cat_matrices(inputVecB, offs, n)
        self.jacobian_K = _cat_matrices(inputVecK, offs, n)
        self.jacobian_Gamma = _cat_tensors(inputVecGamma)
        self.jacobian_Lambda = _cat_tensors(inputVecLambda)
        self.jacobian_Psi = _cat_tensors(inputVecPsi)
        self.jacobian_forcePattern = np.concatenate(
            [np.asarray(f, dtype=np.float64).reshape(-1) for f in inputVecForcePattern])
        self.jacobian_initialConditions = np.concatenate(
            [np.asarray(q, dtype=np.float64).reshape(-1) for q in inputVecInitialConditions])

    def __systemPreprocessing(self):
        """``PASIf.py:607-618``: ``D = -M^{-1}``; B, K, Gamma, Lambda scaled by
        D (row / first index); force pattern times ``-D``."""
        self.system_M = _neg_inverse_diag(self.system_M).tocsc()
        self.system_B = coo_matrix(self.system_M.dot(self.system_B))
        self.system_K = coo_matrix(self.system_M.dot(self.system_K))
        D = self.system_M.diagonal()
        self.system_Gamma.multiplyByDiagMatrix(D)
        self.system_Lambda.multiplyByDiagMatrix(D)
        self.system_forcePattern = self.system_forcePattern * (-1.0 * D)

    def __jacobianPreprocessing(self):
        """``PASIf.py:620-633`` (the adjoint force pattern stays unscaled)."""
        self.jacobian_M = _neg_inverse_diag(self.jacobian_M).tocsc()
        self.jacobian_B = coo_matrix(self.jacobian_M.dot(self.jacobian_B))
        self.jacobian_K = coo_matrix(self.jacobian_M.dot(self.jacobian_K))
        D = self.jacobian_M.diagonal()
        self.jacobian_Gamma.multiplyByDiagMatrix(D)
        self.jacobian_Lambda.multiplyByDiagMatrix(D)
        self.jacobian_Psi.multiplyByDiagMatrix(D)
