"""torch.autograd.Function over brp_adjoint.

``focal_objective(amp, phase, fg0, make_sim, box, n_steps, window)`` returns
J = mean p0^2 over the focal box and the last `window` steps as a 0-d FP64
tensor; its backward hands out dJ/damp, dJ/dphase (per source element) and
dJ/dfg0 (the initial void-fraction map, logical cells [0, N+1] x [0, M+1])
as computed by the library's discrete adjoint in the same call.  The
forward of the Function runs the whole checkpointed forward + reverse sweep
(one brp_adjoint call); backward only scales the stored gradients.

``make_sim(amp, phase)`` must return a fresh, prepared ``Simulation`` with
that source; fg0 (or None) then replaces its initial void-fraction map.
"""
import numpy as np
import torch


class FocalObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, amp, phase, fg0, make_sim, box, n_steps, window, checkpoint_every):
        sim = make_sim(amp.detach().cpu().double().numpy(), phase.detach().cpu().double().numpy())
        try:
            if fg0 is not None:
                sim.set_field("fg", fg0.detach().cpu().double().numpy())
            r = sim.adjoint(box, n_steps, window, checkpoint_every, n_elem=amp.numel())
        finally:
            sim.close()
        # gradients live where (and as what) their inputs do
        ga = torch.from_numpy(np.array(r["grad_amp"])).to(device=amp.device, dtype=amp.dtype)
        gp = torch.from_numpy(np.array(r["grad_phase"])).to(device=phase.device,
                                                             dtype=phase.dtype)
        gf = (torch.from_numpy(np.array(r["grad_fg0"])).to(device=fg0.device, dtype=fg0.dtype)
              if fg0 is not None else None)
        ctx.save_for_backward(ga, gp, *([gf] if gf is not None else []))
        ctx.has_fg = gf is not None
        ctx.stats = {k: r[k] for k in ("forward_ms", "reverse_ms", "memory_bytes",
                                       "n_checkpoints", "segment")}
        return torch.tensor(r["J"], dtype=torch.float64, device=amp.device)

    @staticmethod
    def backward(ctx, gJ):
        saved = ctx.saved_tensors
        ga, gp = saved[0], saved[1]
        gf = saved[2] * gJ.to(saved[2].device) if ctx.has_fg else None
        return (gJ.to(ga.device) * ga, gJ.to(gp.device) * gp, gf, None, None, None, None,
                None)


def focal_objective(amp, phase, fg0, make_sim, box, n_steps, window, checkpoint_every=0):
    return FocalObjective.apply(amp, phase, fg0, make_sim, box, n_steps, window,
                                checkpoint_every)
