"""torch.autograd binding of the adjoint sweep.

``probe_loss(sim, alpha, aex, b1, amp)`` returns the probe spectral loss of
an FM simulation as a differentiable torch scalar of the damping, exchange
constant and B1 of one material and of a per-cell drive-amplitude map on one
layer.  The forward pass runs the native forward with checkpoints
(gf_adjoint_forward), the backward pass the native reverse sweep
(gf_adjoint_backward); nothing is differentiated by torch itself.  With
``precision=32`` the simulation runs the FP32 variant of both sweeps.
"""
from __future__ import annotations

import numpy as np
import torch

from . import FIELD_HDRIVE, FIELD_M, Context, Problem


class ProbeLossSim:
    """One FM context set up for repeated loss evaluations from the same
    initial state at t = 0.  The shaped drive profile is
    drive_vector * amp(x, y) on layer k = drive_layer and zero elsewhere."""

    def __init__(self, problem: Problem, m0, material: int, probe_lo, probe_hi, freq: float, nsteps: int,
                 drive_vector, drive_layer: int, static_fields=None, checkpoint_every: int = 0,
                 precision: int = 64, resident_mib: int = 0):
        self.problem = problem
        self.ctx = Context(problem)
        self.m0 = np.ascontiguousarray(m0, dtype=np.float64)
        self.material = int(material)
        self.base = type(problem.materials[material - 1]).from_buffer_copy(problem.materials[material - 1])
        self.probe = (tuple(probe_lo), tuple(probe_hi))
        self.freq, self.nsteps, self.every = float(freq), int(nsteps), int(checkpoint_every)
        self.precision = int(precision)
        self.resident_mib = int(resident_mib)
        self.vec = np.asarray(drive_vector, dtype=np.float64)
        self.layer = int(drive_layer)
        for fid, v in (static_fields or {}).items():
            self.ctx.upload(fid, v)

    def loss(self, alpha: float, aex: float, b1: float, amp: np.ndarray) -> float:
        mat = type(self.base).from_buffer_copy(self.base)
        mat.FM_damping, mat.Aex, mat.B1 = float(alpha), float(aex), float(b1)
        self.ctx.set_material(self.material, mat)
        prof = np.zeros((3,) + self.ctx.shape)
        for c in range(3):
            prof[c, :, :, self.layer] = self.vec[c] * amp
        self.ctx.upload(FIELD_HDRIVE, prof)
        self.ctx.upload(FIELD_M, self.m0)
        self.ctx.time = 0.0
        return self.ctx.adjoint_forward(self.material, self.probe[0], self.probe[1], self.freq, self.nsteps,
                                        self.every, precision=self.precision, resident_mib=self.resident_mib)

    def grads(self):
        """(dL/dalpha, dL/dAex, dL/dB1, dL/damp) of the last loss."""
        g, ghd = self.ctx.adjoint_backward(drive=True)
        damp = sum(self.vec[c] * ghd[c, :, :, self.layer] for c in range(3))
        return g[0], g[1], g[2], damp

    def close(self) -> None:
        self.ctx.close()


class _ProbeLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, alpha, aex, b1, amp, sim):
        L = sim.loss(float(alpha), float(aex), float(b1), amp.detach().cpu().numpy().astype(np.float64))
        fctx.sim = sim
        fctx.amp_meta = (amp.dtype, amp.device)
        return torch.tensor(L, dtype=torch.float64)

    @staticmethod
    def backward(fctx, gout):
        ga, gA, gB, gamp = fctx.sim.grads()
        s = float(gout)
        dtype, device = fctx.amp_meta
        return (torch.tensor(s * ga, dtype=torch.float64), torch.tensor(s * gA, dtype=torch.float64),
                torch.tensor(s * gB, dtype=torch.float64), torch.from_numpy(s * gamp).to(dtype=dtype, device=device),
                None)


def probe_loss(sim: ProbeLossSim, alpha, aex, b1, amp):
    """Differentiable probe spectral loss (see the module docstring)."""
    return _ProbeLoss.apply(alpha, aex, b1, amp, sim)
