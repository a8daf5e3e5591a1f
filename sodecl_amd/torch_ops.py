"""PyTorch front end: device-resident integration with a discrete adjoint.

    integ = Integrator("kuramoto.cl", solver=2, orbits=16384, neq=64, npar=65,
                       dt=0.01, tspan=20.0, ksteps=10)
    traj = integ(y0, params)           # [nsnap][neq][orbits], float64, cuda
    loss = f(traj); loss.backward()    # y0.grad [orbits][neq], params.grad [orbits][npar]

The forward pass runs the MI355X integrator (libsodecl_amd.so) on the
tensors' memory and writes a checkpoint every ckpt_every steps (default
ceil(sqrt(steps))); backward recomputes each checkpoint segment and sweeps
it backwards (sodecl_adjoint_device): the exact transpose of the Euler /
RK4 step, or for implicit Euler / midpoint the derivative of each step's
implicit equation at its converged Newton solution.  PyTorch provides only
device memory and the stream; all arithmetic is in the HIP kernels.
"""
import os

import torch

from . import sodecl_interface

_SYSTEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "systems")


def _resolve(path):
    if os.path.exists(path):
        return path
    alt = os.path.join(_SYSTEMS, os.path.basename(path))
    return alt if os.path.exists(alt) else path


class _Integrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, params, integ):
        ctx.integ = integ
        if integ.store_stages:
            traj = integ._forward_stored(y0, params)
            ctx.generation = integ._generation
            ctx.save_for_backward(params)
            return traj
        traj, ckpt_t, ckpt_y = integ._forward(y0, params, checkpoints=True)
        ctx.save_for_backward(params, ckpt_t, ckpt_y)
        return traj

    @staticmethod
    def backward(ctx, grad_traj):
        if ctx.integ.store_stages:
            (params,) = ctx.saved_tensors
            gy, gp = ctx.integ._adjoint_stored(params, grad_traj, ctx.generation)
            return gy, gp, None
        params, ckpt_t, ckpt_y = ctx.saved_tensors
        gy, gp = ctx.integ._adjoint(params, ckpt_t, ckpt_y, grad_traj)
        return gy, gp, None


class Integrator:
    """One compiled integrator configuration bound to cuda:<device>."""

    def __init__(self, system, solver, orbits, neq, npar, dt, tspan, ksteps, nnoi=0, device=0, ckpt_every=0,
                 store_stages=False, **options):
        """store_stages: instead of checkpoints, the forward pass keeps every
        step's stage vectors (Euler/RK4; orbits x steps x (stages x neq + 1)
        doubles of device memory) and the backward pass needs no
        recomputation; one forward pass per Integrator may await its
        backward pass at a time."""
        if "schedule" not in options and solver in (1, 2):
            options["schedule"] = "group"
        self.session = sodecl_interface.Session(_resolve(system), solver, orbits, neq, npar, nnoi, dt, tspan, ksteps,
                                                device, options)
        self.orbits, self.neq, self.npar, self.nnoi = orbits, neq, npar, nnoi
        self.device = torch.device("cuda", device)
        self._ready = False
        self.ckpt_every = ckpt_every
        self._work = None
        self.store_stages = store_stages
        self._stored = None
        self._generation = 0

    @property
    def nsnap(self):
        return self.session.nsnap

    def _setup(self, y0, params):
        if not self._ready:
            self.session.setup(y0.detach().cpu().numpy(), params.detach().cpu().numpy())
            if not self.ckpt_every:
                self.ckpt_every = self.session.default_checkpoint_interval()
            self._ready = True

    def _check(self, y0, params):
        for name, t, shape in (("y0", y0, (self.orbits, self.neq)), ("params", params, (self.orbits, self.npar))):
            if t.dtype != torch.float64 or t.device != self.device or tuple(t.shape) != shape:
                raise ValueError(f"{name} must be a float64 tensor of shape {shape} on {self.device}")

    def _forward(self, y0, params, checkpoints=False, counters=None):
        self._check(y0, params)
        self._setup(y0, params)
        y = y0.detach().contiguous().clone()
        p = params.detach().contiguous()
        t = torch.zeros(self.orbits, dtype=torch.float64, device=self.device)
        traj = torch.empty((self.nsnap, self.neq, self.orbits), dtype=torch.float64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        cnt = 0 if counters is None else counters.data_ptr()
        if not checkpoints:
            self.session.run_device(t.data_ptr(), y.data_ptr(), p.data_ptr(), cnt, traj.data_ptr(), stream)
            return traj, None, None
        n = self.session.checkpoint_count(self.ckpt_every)
        ckpt_t = torch.empty((n, self.orbits), dtype=torch.float64, device=self.device)
        ckpt_y = torch.empty((n, self.neq, self.orbits), dtype=torch.float64, device=self.device)
        self.session.run_device_checkpointed(t.data_ptr(), y.data_ptr(), p.data_ptr(), cnt, traj.data_ptr(),
                                             ckpt_t.data_ptr(), ckpt_y.data_ptr(), self.ckpt_every, stream)
        return traj, ckpt_t, ckpt_y

    def _forward_stored(self, y0, params):
        self._check(y0, params)
        self._setup(y0, params)
        need = self.session.stored_workspace_size()
        if need == 0:
            raise RuntimeError("stored stage vectors need Euler or RK4 under the group schedule")
        if self._stored is None or self._stored.numel() < need:
            self._stored = torch.empty(need, dtype=torch.float64, device=self.device)
        y = y0.detach().contiguous().clone()
        p = params.detach().contiguous()
        t = torch.zeros(self.orbits, dtype=torch.float64, device=self.device)
        traj = torch.empty((self.nsnap, self.neq, self.orbits), dtype=torch.float64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.session.run_device_stored(t.data_ptr(), y.data_ptr(), p.data_ptr(), traj.data_ptr(),
                                       self._stored.data_ptr(), stream)
        self._generation += 1
        return traj

    def _adjoint_stored(self, params, grad_traj, generation):
        if generation != self._generation:
            raise RuntimeError("the stored stage vectors belong to a later forward pass of this Integrator")
        cot = grad_traj.contiguous()
        gy = torch.empty((self.orbits, self.neq), dtype=torch.float64, device=self.device)
        gp = torch.empty((self.orbits, self.npar), dtype=torch.float64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.session.adjoint_device_stored(params.detach().contiguous().data_ptr(), cot.data_ptr(), gy.data_ptr(),
                                           gp.data_ptr(), self._stored.data_ptr(), stream)
        return gy, gp

    def _adjoint(self, params, ckpt_t, ckpt_y, grad_traj):
        need = self.session.adjoint_workspace_size(self.ckpt_every)
        if need == 0:
            raise RuntimeError("this configuration has no adjoint (Euler/RK4: group schedule and "
                               "sode_system_vjp_group; implicit Euler/midpoint: lane schedule, calc_jacobian and "
                               "calc_jacobian_params)")
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.float64, device=self.device)
        cot = grad_traj.contiguous()
        gy = torch.empty((self.orbits, self.neq), dtype=torch.float64, device=self.device)
        gp = torch.empty((self.orbits, self.npar), dtype=torch.float64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.session.adjoint_device(ckpt_t.data_ptr(), ckpt_y.data_ptr(), params.detach().contiguous().data_ptr(),
                                    cot.data_ptr(), gy.data_ptr(), gp.data_ptr(), self._work.data_ptr(),
                                    self.ckpt_every, stream)
        return gy, gp

    def __call__(self, y0, params):
        """Trajectory [nsnap][neq][orbits]; differentiable w.r.t. y0 and params."""
        if torch.is_grad_enabled() and (y0.requires_grad or params.requires_grad):
            return _Integrate.apply(y0, params, self)
        return self._forward(y0, params)[0]
