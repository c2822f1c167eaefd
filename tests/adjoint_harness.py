"""torch.autograd.Function over one step of the HIP stepper (test harness only).

forward(x, v, g) runs lsmps_stepper_step with LSMPS_STEP_RECORD on the given
positions (3, N), velocities (3, N) and body force (3,); backward(dL/dx_new,
dL/dv_new) calls lsmps_adjoint_step and returns dL/dx, dL/dv and dL/dg.  The
neighbor lists and labels are those the recorded step built; the adjoint holds
them fixed.
"""
import numpy as np
import torch


class StepFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, v, g, stepper, state, config, abi):
        f = state.fluid
        with torch.no_grad():
            for k, key in enumerate(("x", "y", "z")):
                f[key].copy_(x[k])
            for k, key in enumerate(("vx", "vy", "vz")):
                f[key].copy_(v[k])
        cfg = config.copy()
        for k in range(3):
            cfg.gravity[k] = float(g[k])
        stepper.set_config(cfg)
        stepper.set_mode(abi.STEP_RECORD)
        ctx.report = stepper.step(raise_on_nonconvergence=False)
        ctx.stepper = stepper
        ctx.g_device = g.device
        x_new = torch.stack([f["x"], f["y"], f["z"]]).clone()
        v_new = torch.stack([f["vx"], f["vy"], f["vz"]]).clone()
        return x_new, v_new

    @staticmethod
    def backward(ctx, grad_x_new, grad_v_new):
        if grad_v_new is None:
            grad_v_new = torch.zeros_like(grad_x_new)
        out = ctx.stepper.adjoint(grad_v_new.contiguous(),
                                  grad_x_new.contiguous() if grad_x_new is not None else None)
        ctx.transpose_report = out["report"]
        g = torch.tensor(np.asarray(out["gravity"]), dtype=torch.float64, device=ctx.g_device)
        return out["position"], out["velocity"], g, None, None, None, None
