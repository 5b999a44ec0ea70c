"""torch.autograd.Function over the leapfrog trajectory and its hand-written adjoint.

``trajectory(x0, v0, m, soft, steps=..., dt=...)`` runs ``steps`` kick-drift
steps of the gravity path (Driver::run, src/Driver.h:183-288, with
Partition::kick/perturb) on the GPU through the C ABI and returns the final
positions and velocities as tensors that autograd differentiates with the
library's reverse sweep (pt_leapfrog_adjoint): gradients with respect to the
initial positions, velocities and masses.  Softening is not differentiated.

Tensors cross the boundary as device pointers on the caller's GPU; the
library runs on its own stream and synchronises it before returning, and
the caller's stream is synchronised before each call.  Import this module
(or torch) before creating any Driver: torch reports no GPU when the HIP
runtime was initialised by the library first.
"""
import ctypes

import torch

from . import _lib
from .gravity import FORCE_DIRECT, FORCE_TREE, PT_FP32C, PT_FP64, Configuration, Driver

_PRECISIONS = {"fp64": PT_FP64, "fp32c": PT_FP32C}
_MODES = {"tree": FORCE_TREE, "direct": FORCE_DIRECT}
_hip = None


def _hip_runtime():
    """The HIP runtime already loaded in this process (torch's and the
    library's resolve to the same soname)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def _copy_d2d(dst, src, nbytes):
    st = _hip_runtime().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3)
    if st != 0:
        raise RuntimeError("hipMemcpy device-to-device failed with %d" % st)


def _as_device(t, device):
    return t.detach().to(device=device, dtype=torch.float64).contiguous()


class LeapfrogFunction(torch.autograd.Function):
    """(x0, v0, m) -> (x_T, v_T) through a Driver's pt_leapfrog_forward; the
    backward is pt_leapfrog_adjoint on the same Driver."""

    @staticmethod
    def forward(ctx, x0, v0, m, soft, driver, steps, dt, theta, precision, mode, every,
                adjoint_precision):
        device = x0.device
        if device.type != "cuda":
            raise ValueError("trajectory tensors must be on a GPU device")
        if not torch.cuda.is_available():
            raise RuntimeError("torch sees no GPU: import torch before the first Driver")
        n = x0.shape[0]
        xs = _as_device(x0, device).t().contiguous()
        vs = _as_device(v0, device).t().contiguous()
        md = _as_device(m, device)
        sd = None if soft is None else _as_device(torch.broadcast_to(soft, (n,)), device)
        torch.cuda.current_stream(device).synchronize()
        driver.load_device(n, xs[0].data_ptr(), xs[1].data_ptr(), xs[2].data_ptr(),
                           md.data_ptr(), None if sd is None else sd.data_ptr(),
                           vs[0].data_ptr(), vs[1].data_ptr(), vs[2].data_ptr())
        driver.set_adjoint_precision(adjoint_precision)
        driver.leapfrog_forward(steps, dt, theta, precision, mode, every)
        view = driver.device_view()
        out = torch.empty((6, n), dtype=torch.float64, device=device)
        for k, p in enumerate((view.x, view.y, view.z, view.vx, view.vy, view.vz)):
            _copy_d2d(out[k].data_ptr(), p, n * 8)
        driver._traj_token = getattr(driver, "_traj_token", 0) + 1
        ctx.driver = driver
        ctx.token = driver._traj_token
        ctx.n = n
        return out[:3].t().contiguous(), out[3:].t().contiguous()

    @staticmethod
    def backward(ctx, gx, gv):
        driver = ctx.driver
        if getattr(driver, "_traj_token", None) != ctx.token:
            raise RuntimeError("the Driver ran another trajectory after this one; its "
                               "checkpoints are gone (use one Driver per live trajectory)")
        n = ctx.n
        device = (gx if gx is not None else gv).device
        lx = None if gx is None else _as_device(gx, device)
        lv = None if gv is None else _as_device(gv, device)
        grad_x = torch.empty((n, 3), dtype=torch.float64, device=device)
        grad_v = torch.empty((n, 3), dtype=torch.float64, device=device)
        grad_m = torch.empty((n,), dtype=torch.float64, device=device)
        torch.cuda.current_stream(device).synchronize()
        driver._check(driver.L.pt_leapfrog_adjoint(
            driver.ctx, None if lx is None else lx.data_ptr(),
            None if lv is None else lv.data_ptr(), grad_x.data_ptr(), grad_v.data_ptr(),
            grad_m.data_ptr(), _lib.PT_DEVICE))
        return grad_x, grad_v, grad_m, None, None, None, None, None, None, None, None, None


def trajectory(x0, v0, m, soft=None, *, steps, dt, theta=0.7, precision="fp64", mode="tree",
               checkpoint_every=0, driver=None, bucket=12, adjoint_precision=None):
    """Final (positions, velocities) after `steps` kick-drift steps; both
    differentiable with respect to x0, v0 and m.  `driver` (a gravity.Driver
    on the tensors' GPU) is created with bucket size `bucket` if not given.
    The backward's arithmetic is `adjoint_precision` ("fp64" or "fp32c"),
    by default the forward's `precision`."""
    if driver is None:
        index = x0.device.index if x0.device.index is not None else 0
        driver = Driver(Configuration(max_particles_per_leaf=bucket), device=index)
    adj = precision if adjoint_precision is None else adjoint_precision
    return LeapfrogFunction.apply(x0, v0, m, soft, driver, int(steps), float(dt), float(theta),
                                  _PRECISIONS[precision], _MODES[mode], int(checkpoint_every),
                                  _PRECISIONS[adj])
