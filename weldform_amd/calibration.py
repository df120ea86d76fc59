"""Johnson-Cook / friction calibration through the adjoint of the step.

A `Window` holds a solver positioned at the start of a window of FP64
Leapfrog steps and a target history of the punch force (a rigid plane's
contact force, or a velocity-BC zone's reaction).  `WindowLoss` is a
torch.autograd.Function over theta = (A, B, n, C, m[, mu]):

    loss = WindowLoss.apply(theta, window)      # runs the window forward
    loss.backward()                             # runs the adjoint sweep

Its forward sets the material (and friction) on the device, puts the state
back to the window's start and runs the window (wf_adj_forward); its
backward runs wf_adj_backward.  With `initial_yield=True` the window starts
at the beginning of the run and every particle of the material starts at
the Johnson-Cook yield stress at zero plastic strain,
Sy0 = A (1 - Th^m) (Source/WeldForm.cpp:912-913); its dependence on A and m
is chained into the gradient.

There is no counterpart in the reference: this is the calibration harness of
the differentiable step.
"""
import numpy as np
import torch

from . import capi

JC_KEYS = ("jc_A", "jc_B", "jc_n", "jc_C", "jc_m")


class Window:
    def __init__(self, solver, material, base_material, target, checkpoint_every=20, kind=capi.ADJ_PLANE,
                 index=0, component=2, friction=False, T=None, particles=None, initial_yield=False):
        self.solver = solver
        self.material = int(material)
        self.base = dict(base_material)
        self.target = np.ascontiguousarray(target, dtype=np.float64)
        self.k = int(checkpoint_every)
        self.kind, self.index, self.component = kind, index, component
        self.friction = bool(friction)
        self.initial_yield = bool(initial_yield)
        self.T = None if T is None else np.asarray(T, dtype=np.float64)
        self.sel = particles  # boolean mask of the material's particles
        self.started = False
        self.pending = None
        self.history = None
        if self.initial_yield and (self.T is None or self.sel is None):
            raise ValueError("initial_yield needs the particle temperatures and the material's particle mask")

    def _th(self):
        b = self.base
        return (self.T - b["jc_Tt"]) / (b["jc_Tm"] - b["jc_Tt"])

    def forward(self, theta):
        theta = np.asarray(theta, dtype=np.float64)
        s = self.solver
        if self.started:
            s.adj_restore()
        m = dict(self.base)
        for key, val in zip(JC_KEYS, theta[:5]):
            m[key] = float(val)
        s.set_material(self.material, **m)
        if self.friction:
            s.set_contact(float(theta[5]), float(theta[5]), 0.0)
        if self.initial_yield:
            sy = s.get("sigmay")
            sy[self.sel] = m["jc_A"] * (1.0 - self._th()[self.sel] ** m["jc_m"])
            s.set("sigmay", sy)
        loss, self.history = s.adj_forward(self.target, checkpoint_every=self.k, kind=self.kind, index=self.index,
                                           component=self.component, material=self.material)
        self.started = True
        self.pending = theta.copy()
        return loss

    def backward(self):
        if self.pending is None:
            raise RuntimeError("Window.backward needs the forward of the same parameters first")
        theta = self.pending
        g = self.solver.adj_backward()
        grad = list(g[:5])
        if self.friction:
            grad.append(g[5] + g[6])
        if self.initial_yield:
            syb = self.solver.adj_field("sigmay")[self.sel]
            th = self._th()[self.sel]
            thm = th ** theta[4]
            grad[0] += float(np.sum(syb * (1.0 - thm)))
            lth = np.where(th > 0, np.log(np.where(th > 0, th, 1.0)), 0.0)
            grad[4] += float(np.sum(syb * (-theta[0] * thm * lth)))
        self.pending = None
        return np.array(grad)


class WindowLoss(torch.autograd.Function):
    """loss(theta) of a Window; theta = (A, B, n, C, m[, mu]) as a 1-D tensor."""

    @staticmethod
    def forward(ctx, theta, window):
        ctx.window = window
        loss = window.forward(theta.detach().cpu().double().numpy())
        return theta.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.window.backward()
        return grad_out * torch.as_tensor(g, dtype=grad_out.dtype, device=grad_out.device), None
