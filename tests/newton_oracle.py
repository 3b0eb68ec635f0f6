"""Oracle-driven assembler for the Newton driver (test infrastructure): the same interface
as ``uasource_amd.newton.GpuAssembler``, with R and K from oracle O2 on the host."""
import numpy as np

EPS = np.finfo(float).eps


class OracleAssembler:
    def __init__(self, o2, CtrlVar, MUA, F0, F1, nthreads=8):
        self.o2, self.cv, self.M, self.F0 = o2, CtrlVar, MUA, F0
        self.mua = MUA
        self.F1 = F1.copy()
        self.mesh = o2.O2Mesh(MUA)
        self.nthreads = nthreads
        self.nn = MUA.Nnodes
        self.x = np.concatenate([F1.ub, F1.vb, F1.h]).astype(float)
        self.d = np.zeros_like(self.x)
        self.R0 = None
        self.n_rk = self.n_r = 0
        self.free = None
        self.shift = None

    def set_shift(self, a0=None, a1=None):
        self.shift = None if a0 is None else (np.asarray(a0, dtype=float), np.asarray(a1, dtype=float))

    def set_state(self, x):
        self.x = np.array(x, dtype=float)

    def set_fixed(self, fixed):
        self.free = None
        if fixed is not None and len(fixed):
            self.free = np.ones(3 * self.nn)
            self.free[fixed] = 0.0

    def _F(self, x):
        F = self.F1.copy()
        n = self.nn
        F.ub, F.vb, F.h = x[:n].copy(), x[n:2 * n].copy(), x[2 * n:].copy()
        return F

    def _assemble(self, x, ronly, zero_fields=False):
        return self.o2.assemble(self.cv, self.M, self.F0, self._F(x), mesh=self.mesh, nthreads=self.nthreads,
                                ronly=ronly, zero_fields=zero_fields)

    def normalisation(self):
        R = self._assemble(self.x, True, True)["R"]
        self.n_r += 1
        return float(R @ R) + 1000 * EPS

    def assemble_rk(self, fp32, fp32_residual):
        assert not fp32, "the oracle is FP64"
        out = self._assemble(self.x, False)
        self.n_rk += 1
        self.R0 = out["R"]
        return out["K"].tocsr(), -self.R0

    def set_step(self, d):
        self.d = np.asarray(d, dtype=float).copy()

    def cost(self, gamma, fp32, N):
        if gamma == 0.0:
            R = self.R0
        else:
            R = self._assemble(self.x + gamma * self.d, True)["R"]
            self.n_r += 1
        if self.shift is not None:
            R = R + self.shift[0] + gamma * self.shift[1]
        if self.free is not None:
            R = R * self.free
        D2 = float(-R @ self.d)
        return float(R @ R) / N, D2 * D2, D2

    def accept(self, gamma):
        self.x = self.x + gamma * self.d

    def state(self):
        n = self.nn
        return self.x[:n].copy(), self.x[n:2 * n].copy(), self.x[2 * n:].copy()


class OracleUvAssembler:
    """The uv (diagnostic) system with R and K from oracle O2 (``o2.assemble_uv``): the
    interface of ``uasource_amd.newton.GpuUvAssembler``. ``F`` carries the upper surface s."""

    def __init__(self, o2, CtrlVar, MUA, F, nthreads=8):
        self.o2, self.cv, self.M = o2, CtrlVar, MUA
        self.mua = MUA
        self.F = F.copy()
        self.mesh = o2.O2Mesh(MUA)
        self.nthreads = nthreads
        self.nn = MUA.Nnodes
        self.x = np.concatenate([F.ub, F.vb]).astype(float)
        self.d = np.zeros_like(self.x)
        self.R0 = None
        self.n_rk = self.n_r = 0
        self.free = None
        self.fixed_mask = np.zeros(2 * self.nn)

    def set_shift(self, a0=None, a1=None):
        assert a0 is None

    def set_state(self, x):
        self.x = np.array(x, dtype=float)

    def set_fixed(self, fixed):
        self.free = None
        self.fixed_mask = np.zeros(2 * self.nn)
        if fixed is not None and len(fixed):
            self.free = np.ones(2 * self.nn)
            self.free[fixed] = 0.0
            self.fixed_mask[fixed] = 1.0

    def _assemble(self, x, ronly):
        F = self.F.copy()
        F.ub, F.vb = x[:self.nn].copy(), x[self.nn:].copy()
        return self.o2.assemble_uv(self.cv, self.M, F, mesh=self.mesh, ronly=ronly, nthreads=self.nthreads)

    def normalisation(self):
        R = self._assemble(self.x * self.fixed_mask, True)["R"]
        self.n_r += 1
        return float(R @ R) + 1000 * EPS

    def assemble_rk(self, fp32, fp32_residual):
        assert not fp32, "the oracle is FP64"
        out = self._assemble(self.x, False)
        self.n_rk += 1
        self.R0 = out["R"]
        return out["K"].tocsr(), -self.R0

    def set_step(self, d):
        self.d = np.asarray(d, dtype=float).copy()

    def cost(self, gamma, fp32, N):
        if gamma == 0.0:
            R = self.R0
        else:
            R = self._assemble(self.x + gamma * self.d, True)["R"]
            self.n_r += 1
        if self.free is not None:
            R = R * self.free
        D2 = float(-R @ self.d)
        return float(R @ R) / N, D2 * D2, D2

    def accept(self, gamma):
        self.x = self.x + gamma * self.d

    def state(self):
        n = self.nn
        return self.x[:n].copy(), self.x[n:].copy(), None
