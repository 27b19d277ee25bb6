"""torch.autograd view of the wave-equation core.

    engine = Engine(geom, shots)
    traces = simulate(vel, engine)          # forward, every shot in one batch
    loss = 0.5 * ((traces - obs) ** 2).sum()
    loss.backward()                         # vel.grad = exact adjoint gradient

The forward is fwi_hip_forward; the backward hands dL/dtraces to
fwi_hip_set_residual and runs fwi_hip_adjoint (the hand-written exact
transpose of the forward, with source-field reconstruction and imaging),
returning dL/dvel.  CUDA (HIP) tensors cross the C ABI as device pointers:
the traces, the residual and the gradient are copied device to device and
never pass through host memory (the model, 4 B per cell, is staged through
the host for its stability check).  CPU tensors go through host memory.  The
result comes back on the device of the input.

When torch tensors live on the GPU, import torch before this module loads
the library: the loader then links the build of the library against
torch's own HIP runtime (libfwi_hip_torch.so), so the process holds one
HIP runtime.
"""
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

    def run_forward(self, vel):
        """vel: a float32 contiguous CPU or GPU tensor; returns the packed traces
        on vel's device and the id of this forward."""
        out = torch.empty(self.ctx.ntrace(), dtype=torch.float32, device=vel.device)
        if vel.is_cuda:
            torch.cuda.current_stream(vel.device).synchronize()   # vel complete before the context's stream reads it
        self.ctx._chk(self.ctx.lib.fwi_hip_set_model(self.ctx.h, _ptr(vel)))
        self.ctx._chk(self.ctx.lib.fwi_hip_forward(self.ctx.h, _ptr(out)))
        self._fwd_id = self._next_id
        self._next_id += 1
        return out, self._fwd_id

    def run_adjoint(self, vel, fwd_id, res):
        """vel, res: float32 contiguous tensors (CPU or GPU); dL/dvel in FP64 on vel's device."""
        if fwd_id != self._fwd_id:
            # another forward ran in between: rebuild the frames of this one
            self.ctx._chk(self.ctx.lib.fwi_hip_set_model(self.ctx.h, _ptr(vel)))
            self.ctx._chk(self.ctx.lib.fwi_hip_forward(self.ctx.h, None))
            self._fwd_id = fwd_id
        if res.is_cuda:
            torch.cuda.current_stream(res.device).synchronize()
        self.ctx._chk(self.ctx.lib.fwi_hip_set_residual(self.ctx.h, _ptr(res)))
        g = self.geom
        grad = torch.empty((g.nz, g.nx), dtype=torch.float64, device=vel.device)
        self.ctx._chk(self.ctx.lib.fwi_hip_adjoint(self.ctx.h, _ptr(grad), None))
        self._fwd_id = None  # the adjoint consumed this forward's final wavefield
        return grad


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


class WaveForward(torch.autograd.Function):
    """traces = F(vel); backward gives (dL/dtraces)^T dF/dvel."""

    @staticmethod
    def forward(ctx, vel, engine):
        v = vel.detach().to(torch.float32).contiguous()
        tr, fid = engine.run_forward(v)
        ctx.engine = engine
        ctx.fid = fid
        ctx.vel = v
        return tr

    @staticmethod
    def backward(ctx, gtr):
        res = gtr.detach().to(device=ctx.vel.device, dtype=torch.float32).contiguous()
        grad = ctx.engine.run_adjoint(ctx.vel, ctx.fid, res)
        return grad.to(torch.float32), None


def simulate(vel, engine):
    """Packed receiver traces of every shot of `engine` for velocity `vel`."""
    return WaveForward.apply(vel, engine)
