"""torch.autograd.Function around the hand-written adjoint (PyTorch is used only here).

    problem = adjoint.UtciGradient(domain, objective_hour=14)
    J = utci_objective(problem, alb, emis, albedo_b, trans, Ta, RH, Ws, radG)
    J.backward()

``alb``/``emis`` are [rows, cols] tensors (host or device), the others scalars or [n_hours]
tensors. The forward runs the batched forward up to the objective hour; the backward runs the
device sweep (the forward of the earlier hours recomputed, then the reverse hours) and returns
the gradients on the devices the inputs live on.
"""
from __future__ import annotations

import numpy as np
import torch

from . import runtime as rt


class _DevicePlane:
    """A torch device tensor seen as a device pointer of the library."""

    def __init__(self, t):
        self.ptr, self.nbytes = t.data_ptr(), t.numel() * t.element_size()


def _map_input(t, shape):
    t = t.detach().to(torch.float32).contiguous()
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f'map of shape {tuple(t.shape)}, expected {tuple(shape)}')
    if t.is_cuda:
        torch.cuda.synchronize(t.device)
        return _DevicePlane(t), t
    return t.numpy(), t


def _hourly(v, n):
    a = np.asarray(v.detach().cpu().numpy() if torch.is_tensor(v) else v, dtype=np.float64)
    return np.broadcast_to(a, (n,)).copy()


class UtciObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, alb, emis, albedo_b, trans, Ta, RH, Ws, radG):
        n = problem.n_hours
        met = problem.met.copy()
        for col, v in ((11, Ta), (10, RH), (9, Ws), (14, radG)):
            met[:, col] = _hourly(v, n)
        a, keep_a = _map_input(alb, problem.pd.shape)
        e, keep_e = _map_input(emis, problem.pd.shape)
        problem.set_parameters(alb=a if not isinstance(a, _DevicePlane) else _as_view(a, problem, 'alb'),
                               emis=e if not isinstance(e, _DevicePlane) else _as_view(e, problem, 'emis'),
                               albedo_b=float(albedo_b), trans=float(trans), met=met)
        del keep_a, keep_e
        ctx.problem = problem
        ctx.devices = (alb.device, emis.device)
        ctx.map_dtypes = (alb.dtype, emis.dtype)
        ctx.hourly = tuple(torch.is_tensor(v) and v.dim() > 0 for v in (Ta, RH, Ws, radG))
        ctx.kinds = tuple((v.dtype, v.device) if torch.is_tensor(v) else (torch.float64, torch.device('cpu'))
                          for v in (albedo_b, trans, Ta, RH, Ws, radG))
        return torch.tensor(problem.objective(), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g = ctx.problem.gradient()
        s = float(gJ)

        def plane(key, device):
            if device.type == 'cuda':
                out = torch.empty(ctx.problem.pd.shape, dtype=torch.float32, device=device)
                torch.cuda.synchronize(device)
                rt.call('sw_memcpy_d2d', out.data_ptr(), ctx.problem.buf['grad_' + key].ptr, out.numel() * 4,
                        ctx.problem.stream.handle)
                ctx.problem.stream.sync()
                return out * s
            return torch.from_numpy(np.ascontiguousarray(g[key], dtype=np.float32)) * s

        need = ctx.needs_input_grad
        out = [None,
               plane('alb', ctx.devices[0]).to(ctx.map_dtypes[0]) if need[1] else None,
               plane('emis', ctx.devices[1]).to(ctx.map_dtypes[1]) if need[2] else None,
               torch.tensor(g['albedo_b'] * s, dtype=ctx.kinds[0][0], device=ctx.kinds[0][1]) if need[3] else None,
               torch.tensor(g['trans'] * s, dtype=ctx.kinds[1][0], device=ctx.kinds[1][1]) if need[4] else None]
        for k, (name, is_vec) in enumerate(zip(('Ta', 'RH', 'Ws', 'radG'), ctx.hourly)):
            if not need[5 + k]:
                out.append(None)
                continue
            v = torch.from_numpy(g[name]) * s
            dt, dev = ctx.kinds[2 + k]
            out.append((v if is_vec else v.sum()).to(dtype=dt, device=dev))
        return tuple(out)


def _as_view(dp, problem, key):
    """Copy a device tensor's plane into the problem's map (device to device)."""
    rt.call('sw_memcpy_d2d', problem.pd.dev[key].ptr, dp.ptr, dp.nbytes, problem.stream.handle)
    return None


def utci_objective(problem, alb, emis, albedo_b, trans, Ta, RH, Ws, radG):
    return UtciObjective.apply(problem, alb, emis, albedo_b, trans, Ta, RH, Ws, radG)
