"""Body of test_gpu_autograd (run in its own process: torch is imported
before the library, so the library linked against torch's HIP runtime is the
one loaded, and CUDA tensors cross the autograd boundary)."""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [os.path.join(ROOT, "fwi2dacoustic-mpi-lbfgs_amd"), os.path.join(ROOT, "oracle")]

from fwihip import lib as L  # noqa: E402
from fwihip.autograd import Engine, simulate  # noqa: E402
from fwihip.workloads import Geometry, Shot, ricker  # noqa: E402
from oracle import Oracle  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def main():
    if not torch.cuda.is_available():
        print("no GPU")
        return
    dev = torch.device("cuda")
    g = Geometry(nx=100, nz=50, npml=10, dx=10.0, dz=10.0, dt=1e-3)
    rng = np.random.default_rng(3)
    vtrue = np.clip(2300 + 200 * rng.standard_normal((g.nz, g.nx)), 1700, 3000).astype(np.float32)
    nt = 260
    shots = []
    for j, sx in enumerate((20, 70)):
        irx = np.arange(0, g.nx, 2, dtype=np.int32)
        shots.append(Shot(nt=nt, isx=sx, isz=2, wav=ricker(nt, g.dt, 20.0), irx=irx,
                          irz=np.full(len(irx), 2, np.int32)))
    eng = Engine(g, shots)
    with torch.no_grad():
        obs = simulate(torch.from_numpy(vtrue).to(dev), eng)
    v0 = torch.from_numpy(vtrue * 0.97).to(dev).requires_grad_(True)
    tr = simulate(v0, eng)
    assert L.load()._name.endswith("libfwi_hip_torch.so"), L.load()._name
    loss = 0.5 * ((tr - obs) ** 2).sum()
    loss.backward()
    grad = v0.grad.detach().cpu().numpy().astype(np.float64)
    # oracle: sum of per-shot FP32 gradients
    o = Oracle(np.float32)
    obs_np = obs.cpu().numpy()
    gsum = np.zeros((g.nz, g.nx))
    off = 0
    for s in shots:
        ob = obs_np[off:off + s.nr * s.nt].reshape(s.nr, s.nt)
        off += s.nr * s.nt
        gsum += o.gradient(g, vtrue * 0.97, s, ob)[2]
    r1 = rel(grad, gsum)
    print(f"autograd grad vs oracle: {r1:.2e}")
    assert r1 < 1e-4
    # a second forward in between: backward must still use its own forward
    v1 = torch.from_numpy(vtrue * 1.02).to(dev).requires_grad_(True)
    tr1 = simulate(v1, eng)
    _ = simulate(torch.from_numpy(vtrue * 0.9).to(dev), eng)
    (0.5 * ((tr1 - obs) ** 2).sum()).backward()
    g1 = v1.grad.detach().cpu().numpy()
    v1b = torch.from_numpy(vtrue * 1.02).to(dev).requires_grad_(True)
    (0.5 * ((simulate(v1b, eng) - obs) ** 2).sum()).backward()
    assert rel(g1, v1b.grad.detach().cpu().numpy()) < 1e-6
    # directional finite difference (central, FP32 forward), smooth direction
    zz, xx = np.mgrid[0:g.nz, 0:g.nx]
    dv = 50.0 * np.sin(np.pi * xx / (g.nx - 1)) * np.sin(np.pi * (zz + 1) / g.nz)
    dv[:, 0] = dv[:, -1] = dv[-1, :] = 0.0
    dvt = torch.from_numpy(dv.astype(np.float32)).to(dev)

    def J(v):
        with torch.no_grad():
            return float((0.5 * ((simulate(v, eng) - obs).double() ** 2).sum()).item())
    gd = float(np.sum(grad * dv))
    vbase = torch.from_numpy(vtrue * 0.97).to(dev)
    J0 = J(vbase)
    errs, rem = [], []
    # Taylor ladder: the first-order remainder falls as h^2 down to the FP32
    # rounding floor of the misfit; the central difference meets the gradient
    for h in (1e-1, 1e-2, 1e-3, 1e-4):
        jp, jm = J(vbase + h * dvt), J(vbase - h * dvt)
        fd = (jp - jm) / (2 * h)
        errs.append(abs(fd - gd) / abs(gd))
        rem.append(abs(jp - J0 - h * gd))
        print(f"h={h:.0e}: remainder {rem[-1]:.3e} fd {fd:.6e} adjoint {gd:.6e} rel {errs[-1]:.2e}")
    assert 25 < rem[0] / rem[1] < 400
    assert min(errs) < 2e-3
    # CPU tensors take the host-memory path through the same calls: the same
    # gradient bit for bit as the device-pointer path above
    vc = torch.from_numpy(vtrue * 0.97).requires_grad_(True)
    (0.5 * ((simulate(vc, eng) - obs.cpu()) ** 2).sum()).backward()
    assert vc.grad.device.type == "cpu" and v0.grad.device.type == dev.type
    np.testing.assert_array_equal(vc.grad.numpy(), v0.grad.detach().cpu().numpy())
    eng.close()
    print("autograd ok")


if __name__ == "__main__":
    main()
