"""torch.autograd.Function over the C ABI: a differentiable smoke rollout.

forward  = ds_simulate (the forward steps, model.cpp:341-349, driven as solver::simulate,
           solver.ixx:53-81);
backward = ds_trajectory_vjp (the hand-written VJP, model.cpp:361-389, driven as solver::vjp,
           solver.ixx:115-150, with a checkpoint every `checkpoint_interval` steps and recompute);
jvp      = ds_trajectory_jvp (forward mode, model.cpp:351-359, as solver::jvp, solver.ixx:83-113,
           in one sweep), for torch.autograd.forward_ad / torch.func.jvp.

Tensors stay on the GPU (zero copy): the C ABI reads and writes them in place. The control is the
same at every step, so its gradient is the sum over steps. Parameters are one float32 tensor
(ambient temperature, density buoyancy, temperature buoyancy, vorticity confinement); their
adjoints are accumulated in FP64 by the library and returned as float32.
"""
from __future__ import annotations

import torch

from .solver import Context

STATE_KEYS = ("density", "temperature", "velocity_x", "velocity_y", "velocity_z")
CONTROL_KEYS = ("density_source", "temperature_source", "acceleration_x", "acceleration_y", "acceleration_z")


def _state(t):
    return {"density": t[0], "temperature": t[1], "velocity": (t[2], t[3], t[4])}


def _control(t):
    return {"density_source": t[0], "temperature_source": t[1], "external_acceleration": (t[2], t[3], t[4])}


def _flat_state(d):
    return [d["density"], d["temperature"], *d["velocity"]]


def _flat_control(d):
    return [d["density_source"], d["temperature_source"], *d["external_acceleration"]]


def _checked(solver: Context, tensors, shapes, what):
    out = []
    for t, shape in zip(tensors, shapes):
        if t.device != solver.device or t.dtype != torch.float32:
            raise ValueError(f"{what}: tensors must be float32 on {solver.device}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{what}: shape {tuple(t.shape)} does not match the grid ({tuple(shape)})")
        out.append(t.detach().contiguous())
    return out


class SmokeRollout(torch.autograd.Function):
    @staticmethod
    def forward(fctx, solver: Context, steps: int, checkpoint_interval: int, params, *fields):
        from .scenes import cell_shape, face_shape

        n = solver.n
        cs = cell_shape(n)
        state = _checked(solver, fields[:5], [cs, cs] + [face_shape(n, a) for a in range(3)], "state")
        control = _checked(solver, fields[5:], [cs] * 5, "control")
        (params,) = _checked(solver, [params], [(4,)], "params")
        current = torch.cuda.current_stream(solver.device)
        solver.stream.wait_stream(current)
        final = solver.allocate_state()
        solver.simulate(_state(state), _control(control), params, steps, final_state=final)
        current.wait_stream(solver.stream)
        outs = _flat_state(final)
        for t in outs:
            t.record_stream(current)
        fctx.solver = solver
        fctx.steps = steps
        fctx.checkpoint_interval = checkpoint_interval
        fctx.save_for_backward(params, *state, *control)
        fctx.save_for_forward(params, *state, *control)
        return tuple(outs)

    @staticmethod
    def jvp(fctx, _solver, _steps, _checkpoint_interval, params_t, *fields_t):
        """Forward mode: ds_trajectory_jvp (solver::jvp, solver.ixx:83-113, in one sweep that runs the
        forward again); a missing tangent is zero."""
        solver = fctx.solver
        saved = fctx.saved_tensors
        params, state, control = saved[0], list(saved[1:6]), list(saved[6:11])
        zero_or = lambda t, like: t.detach().contiguous() if t is not None else torch.zeros_like(like)
        state_t = [zero_or(t, s) for t, s in zip(fields_t[:5], state)]
        control_t = [zero_or(t, k) for t, k in zip(fields_t[5:], control)]
        params_t = zero_or(params_t, params)
        current = torch.cuda.current_stream(solver.device)
        solver.stream.wait_stream(current)
        _, final_t, _ = solver.trajectory_jvp(_state(state), _control(control), params, fctx.steps, _state(state_t),
                                              _control(control_t), params_t)
        current.wait_stream(solver.stream)
        outs = _flat_state(final_t)
        for t in outs:
            t.record_stream(current)
        return tuple(outs)

    @staticmethod
    def backward(fctx, *grads):
        solver = fctx.solver
        saved = fctx.saved_tensors
        params, state, control = saved[0], list(saved[1:6]), list(saved[6:11])
        current = torch.cuda.current_stream(solver.device)
        final_adj = _state([g.contiguous() if g is not None else torch.zeros_like(s) for g, s in zip(grads, state)])
        solver.stream.wait_stream(current)
        init_adj = solver.allocate_state_adjoint()
        want_control = any(fctx.needs_input_grad[4 + 5:])
        ctl_adj = solver.allocate_control_adjoint() if want_control else None
        par_adj = solver.allocate_parameter_adjoint() if fctx.needs_input_grad[3] else None
        solver.trajectory_vjp(_state(state), _control(control), params, fctx.steps, fctx.checkpoint_interval, final_adj,
                              initial_adj=init_adj, control_adj=ctl_adj, param_adj=par_adj)
        current.wait_stream(solver.stream)
        g_state = _flat_state(init_adj)
        g_control = _flat_control(ctl_adj) if ctl_adj is not None else [None] * 5
        g_params = par_adj.to(torch.float32) if par_adj is not None else None
        for t in g_state + [t for t in g_control if t is not None] + ([g_params] if g_params is not None else []):
            t.record_stream(current)
        return (None, None, None, g_params, *g_state, *g_control)


def rollout(solver: Context, params: torch.Tensor, state: dict, control: dict, steps: int, checkpoint_interval: int = 8):
    """Differentiable rollout: returns the final state {"density", "temperature", "velocity": (x, y, z)}.

    state / control: dicts in the layout of ds_smoke.solver.Context (torch tensors on the
    context's device); gradients flow to every tensor that requires them and to params."""
    outs = SmokeRollout.apply(solver, int(steps), int(checkpoint_interval), params, *_flat_state(state), *_flat_control(control))
    return _state(list(outs))
