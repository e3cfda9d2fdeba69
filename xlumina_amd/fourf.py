"""Batched 4f-system optimisation: B candidate set-ups per launch on one GPU.

Each candidate b holds two phase masks phi_b = (phi_b1, phi_b2) (n x n,
float32).  Its optical table is the reference's dual-SLM 4f system
(experiments/four_f_optical_table.py:36-83) with CZT propagators:

    E_in --CZT(f)--> a1 --SLM(phi_b1)--CZT(2f)--> a2 --SLM(phi_b2)--CZT(f)--> a3

and its loss is MSE_Intensity(a3, T) = sum_p (|a3|^2 - |T|^2)^2 / n^2
(xlumina/loss_functions.py:137-149).  The phases are trained by Adam / AdamW
(optax.adam / optax.adamw, experiments/four_f_optimizer.py:112).

Device work per step, every launch covering all B candidates:
  1. CZT(2f) of the shared a1 with SLM phi_b1 fused into the row pass  -> a2
  2. CZT(f) of a2 with SLM phi_b2 fused                                  -> a3
  3. loss and cotangent g3 = 4 (|a3|^2 - T) a3 / n^2 (in place)
  4. adjoint CZT(f) of g3 -> cotangent of the phi_b2-masked field
  5. SLM adjoint: d loss / d phi_b2 and the cotangent of a2, with the Adam
     update of phi_b2 fused in (the phase gradient never reaches HBM)
  6. adjoint CZT(2f) -> cotangent of the phi_b1-masked field
  7. SLM adjoint + Adam update of phi_b1
  8. the Adam step counter advances.
a1 does not depend on the phases and is computed once; a2 is the only
activation kept for the backward pass.  `steps_per_launch` steps are captured
into one HIP graph, so a launch is one graph replay.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import ops
from .ops import MaskSpec
from .toolbox import space


def gaussian_field(x, y, wavelength: float, w0: float) -> np.ndarray:
    """LightSource.gaussian_beam at its waist, E0 = 1 (wave_optics.py:468-523)."""
    X, Y = np.meshgrid(x, y)
    k = 2 * np.pi / wavelength
    phase = np.exp(-1j * (k * X ** 2 / (2 * 1e12) + k * Y ** 2 / (2 * 1e12)))
    return np.exp(-X ** 2 / w0 ** 2 - Y ** 2 / w0 ** 2) * phase


def seeded_phases(batch: int, n: int, seed: int = 5) -> np.ndarray:
    """rng(seed).uniform(-pi, pi, (batch, 2, n, n)) as float32, drawn per candidate."""
    rng = np.random.default_rng(seed)
    out = np.empty((batch, 2, n, n), dtype=np.float32)
    for b in range(batch):
        out[b] = rng.uniform(-np.pi, np.pi, (2, n, n))
    return out


def seeded_target(n: int, seed: int = 6) -> np.ndarray:
    """Random binary amplitude image: |T|^2 = T."""
    return (np.random.default_rng(seed).random((n, n)) > 0.5).astype(np.float32)


class FourFOptimizer:
    def __init__(self, device=None, batch: int = 64, n: int = 1024, half_width: float = 1200.0,
                 wavelength: float = 0.65, f: float = 25e3, w0: float = 1200.0, lr: float = 0.05,
                 b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0,
                 phases: Optional[np.ndarray] = None, target: Optional[np.ndarray] = None,
                 steps_per_launch: int = 100, use_graph: bool = True):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        N.ensure_device(dev)
        self.device, self.batch, self.n = dev, batch, n
        self.x, self.y = space(half_width, n)
        self.wavelength, self.f = wavelength, f
        self.lr, self.b1, self.b2, self.eps, self.wd = lr, b1, b2, eps, weight_decay
        self.g_f = ops.czt_geometry(self.x, self.y, self.x, self.y, f, wavelength)
        self.g_2f = ops.czt_geometry(self.x, self.y, self.x, self.y, 2 * f, wavelength)
        self.e_in = torch.from_numpy(gaussian_field(self.x, self.y, wavelength, w0).astype(np.complex64)).to(dev)
        self.a1 = ops.czt(self.e_in[None], self.g_f, form="bluestein")     # (1, n, n), shared
        ph = seeded_phases(batch, n) if phases is None else np.asarray(phases, dtype=np.float32)
        if ph.shape != (batch, 2, n, n):
            raise ValueError(f"phases must be {(batch, 2, n, n)}")
        self.phi = torch.from_numpy(ph).to(dev)
        tg = seeded_target(n) if target is None else np.asarray(target, dtype=np.float32)
        self.target = torch.from_numpy(np.ascontiguousarray(tg)).to(dev)
        self.m = torch.zeros_like(self.phi)
        self.v = torch.zeros_like(self.phi)
        self.grad = torch.zeros_like(self.phi)
        self.count = torch.zeros(1, dtype=torch.int32, device=dev)
        shape = (batch, 1, n, n)
        self.a2 = torch.empty(shape, dtype=torch.complex64, device=dev)
        self.buf_a = torch.empty(shape, dtype=torch.complex64, device=dev)   # a3 -> g3 -> cotangent of a2
        self.buf_b = torch.empty(shape, dtype=torch.complex64, device=dev)   # masked-field cotangents
        ws = max(ops.czt_workspace_bytes(self.g_f, batch=batch), ops.czt_workspace_bytes(self.g_f, adjoint=True,
                                                                                          batch=batch))
        self.ws = torch.empty(ws, dtype=torch.uint8, device=dev)
        self.mse_ws = torch.empty(int(N.load().xl_mse_workspace_bytes(n * n, batch)), dtype=torch.uint8, device=dev)
        self.steps_per_launch = steps_per_launch
        self.losses = torch.zeros((steps_per_launch, batch), dtype=torch.float32, device=dev)
        self.use_graph = use_graph
        self.graph = None
        self.stream = torch.cuda.Stream(device=dev)

    # -- one step -------------------------------------------------------------
    def forward_backward(self, loss_out: torch.Tensor) -> None:
        """Steps 1-7: loss_out (B,) and self.grad (B, 2, n, n)."""
        B, n = self.batch, self.n
        ph1, ph2 = self.phi[:, 0], self.phi[:, 1]
        ops.czt(self.a1, self.g_2f, masks=[MaskSpec(N.XL_MASK_SLM, ph1)], batch=B, out=self.a2,
                workspace=self.ws)
        ops.czt(self.a2, self.g_f, masks=[MaskSpec(N.XL_MASK_SLM, ph2)], batch=B, out=self.buf_a,
                workspace=self.ws)
        fa = self.buf_a.view(B, n, n)
        ops.mse_intensity(fa, self.target, grad=fa, loss=loss_out, workspace=self.mse_ws)
        ops.czt(self.buf_a, self.g_f, adjoint=True, batch=B, out=self.buf_b, workspace=self.ws)
        ops.pre_backward_batched(self.a2, self.buf_b, self.x, self.y, B, masks=[MaskSpec(N.XL_MASK_SLM, ph2)],
                                 need_src=True, out=self.buf_a, need_params=((True, False), (False, False)),
                                 gparams=[[self.grad[:, 1], None], [None, None]])
        ops.czt(self.buf_a, self.g_2f, adjoint=True, batch=B, out=self.buf_b, workspace=self.ws)
        ops.pre_backward_batched(self.a1, self.buf_b, self.x, self.y, B, masks=[MaskSpec(N.XL_MASK_SLM, ph1)],
                                 need_src=False, need_params=((True, False), (False, False)),
                                 gparams=[[self.grad[:, 0], None], [None, None]])

    def forward_intensity(self) -> torch.Tensor:
        """Detector intensities (B, n, n) of the current phases, through the
        same batched launches as a training step's forward half."""
        B, n = self.batch, self.n
        ph1, ph2 = self.phi[:, 0], self.phi[:, 1]
        a2 = torch.empty_like(self.a2)
        a3 = torch.empty_like(self.buf_a)
        ops.czt(self.a1, self.g_2f, masks=[MaskSpec(N.XL_MASK_SLM, ph1)], batch=B, out=a2, workspace=self.ws)
        ops.czt(a2, self.g_f, masks=[MaskSpec(N.XL_MASK_SLM, ph2)], batch=B, out=a3, workspace=self.ws)
        a3 = a3.view(B, n, n)
        return a3.real ** 2 + a3.imag ** 2

    def step(self, loss_out: torch.Tensor) -> None:
        """One training step.  The mask adjoints run fused with the Adam update
        of their phases (xl_slm_backward_adam), so no gradient array is written:
        phi_b2 is updated as soon as its cotangent exists (nothing later in the
        step reads it), then phi_b1; the step counter advances once at the end.
        `self.grad` is filled only by forward_backward()."""
        B, n = self.batch, self.n
        ph1, ph2 = self.phi[:, 0], self.phi[:, 1]
        ops.czt(self.a1, self.g_2f, masks=[MaskSpec(N.XL_MASK_SLM, ph1)], batch=B, out=self.a2,
                workspace=self.ws)
        ops.czt(self.a2, self.g_f, masks=[MaskSpec(N.XL_MASK_SLM, ph2)], batch=B, out=self.buf_a,
                workspace=self.ws)
        fa = self.buf_a.view(B, n, n)
        ops.mse_intensity(fa, self.target, grad=fa, loss=loss_out, workspace=self.mse_ws)
        ops.czt(self.buf_a, self.g_f, adjoint=True, batch=B, out=self.buf_b, workspace=self.ws)
        hp = (self.lr, self.b1, self.b2, self.eps, self.wd)
        ops.slm_backward_adam(self.a2, self.buf_b, self.x, self.y, B, ph2, self.m[:, 1], self.v[:, 1], self.count,
                              *hp, out=self.buf_a)
        ops.czt(self.buf_a, self.g_2f, adjoint=True, batch=B, out=self.buf_b, workspace=self.ws)
        ops.slm_backward_adam(self.a1, self.buf_b, self.x, self.y, B, ph1, self.m[:, 0], self.v[:, 0], self.count,
                              *hp)
        ops.adam_tick(self.count)

    def step_unfused(self, loss_out: torch.Tensor) -> None:
        """The same step as forward_backward() + one xl_adam_step over all phases."""
        self.forward_backward(loss_out)
        ops.adam_step(self.phi, self.grad, self.m, self.v, self.count, self.lr, self.b1, self.b2, self.eps,
                      self.wd)

    def _launch_body(self):
        for s in range(self.steps_per_launch):
            self.step(self.losses[s])

    def capture(self) -> None:
        """Captures steps_per_launch steps into one HIP graph (after one eager
        warm-up launch whose updates are undone)."""
        saved = [t.clone() for t in (self.phi, self.m, self.v, self.count)]
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            self.step(self.losses[0])
        self.stream.synchronize()
        for t, s in zip((self.phi, self.m, self.v, self.count), saved):
            t.copy_(s)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self._launch_body()
        self.stream.synchronize()

    def launch(self) -> torch.Tensor:
        """Runs steps_per_launch optimisation steps; returns the (steps, B)
        loss history of this launch (the loss before each step's update)."""
        if self.use_graph:
            if self.graph is None:
                self.capture()
            self.stream.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(self.stream):
                self.graph.replay()
        else:
            self.stream.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(self.stream):
                self._launch_body()
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
        return self.losses

    def candidate_steps_per_launch(self) -> int:
        return self.batch * self.steps_per_launch
