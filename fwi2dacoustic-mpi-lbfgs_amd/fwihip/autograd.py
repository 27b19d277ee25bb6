"""torch.autograd view of the wave-equation core.

    engine = Engine(geom, shots)
    traces = simulate(vel, engine)          # forward, every shot in one batch
    loss = 0.5 * ((traces - obs) ** 2).sum()
    loss.backward()                         # vel.grad = exact adjoint gradient

The forward is fwi_hip_forward; the backward hands dL/dtraces to
fwi_hip_set_residual and runs fwi_hip_adjoint (the hand-written exact
transpose of the forward, with source-field reconstruction and imaging),
returning dL/dvel.  Tensors cross the C ABI through host memory, so `vel`
and the incoming gradient may live on any device; the result comes back on
the device of the input.

When torch tensors live on the GPU, import torch before this module loads
the library: the loader then links the build of the library against
torch's own HIP runtime (libfwi_hip_torch.so), so the process holds one
HIP runtime.
"""
import numpy as np
import torch

from . import lib as L


class Engine:
    """A context holding one batch of shots; reused across forward/backward."""

    def __init__(self, geom, shots, tile_z=0, shots_per_group=0):
        self.geom = geom
        self.shots = list(shots)
        nt = max(s.nt for s in self.shots)
        nr = max(s.nr for s in self.shots)
        self.ctx = L.Context(geom, max_shots=len(self.shots), max_nt=nt, max_nr=nr, tile_z=tile_z,
                             shots_per_group=shots_per_group)
        self.ctx.set_shots(self.shots)
        self._fwd_id = 0      # id of the forward whose frames the context holds
        self._next_id = 1

    def close(self):
        self.ctx.close()

    def run_forward(self, vel_np):
        self.ctx.set_model(vel_np)
        tr = self.ctx.forward()
        self._fwd_id = self._next_id
        self._next_id += 1
        return tr, self._fwd_id

    def run_adjoint(self, vel_np, fwd_id, res_np):
        if fwd_id != self._fwd_id:
            # another forward ran in between: rebuild the frames of this one
            self.ctx.set_model(vel_np)
            self.ctx.forward(want_traces=False)
            self._fwd_id = fwd_id
        self.ctx.set_residual(res_np)
        grad, _ = self.ctx.adjoint()
        self._fwd_id = None  # the adjoint consumed this forward's final wavefield
        return grad


class WaveForward(torch.autograd.Function):
    """traces = F(vel); backward gives (dL/dtraces)^T dF/dvel."""

    @staticmethod
    def forward(ctx, vel, engine):
        vel_np = np.ascontiguousarray(vel.detach().to("cpu", torch.float32).numpy())
        tr, fid = engine.run_forward(vel_np)
        ctx.engine = engine
        ctx.fid = fid
        ctx.vel_np = vel_np
        ctx.device = vel.device
        return torch.from_numpy(tr).to(vel.device)

    @staticmethod
    def backward(ctx, gtr):
        res = np.ascontiguousarray(gtr.detach().to("cpu", torch.float32).numpy())
        grad = ctx.engine.run_adjoint(ctx.vel_np, ctx.fid, res)
        return torch.from_numpy(grad.astype(np.float32)).to(ctx.device), None


def simulate(vel, engine):
    """Packed receiver traces of every shot of `engine` for velocity `vel`."""
    return WaveForward.apply(vel, engine)
