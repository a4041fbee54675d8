"""The torch.autograd.Function harness over brp_adjoint, with a stub
simulation (CPU only): gradients come back on their inputs' devices and in
their dtypes, scaled by the incoming gradient."""
import importlib

import numpy as np
import torch

ag = importlib.import_module("cuda-rp_amd.autograd")


class _StubSim:
    """Stands in for Simulation: records what it was given and returns fixed
    adjoint results (J, per-element gradients, a gradient map)."""

    def __init__(self, amp, phase):
        self.amp, self.phase, self.fg = amp, phase, None

    def set_field(self, name, arr):
        assert name == "fg"
        self.fg = arr

    def adjoint(self, box, n_steps, window, checkpoint_every, n_elem):
        assert n_elem == len(self.amp)
        return {"J": 2.5, "grad_amp": 3.0 * np.arange(n_elem, dtype=np.float64),
                "grad_phase": -np.ones(n_elem), "grad_fg0": np.full(self.fg.shape, 7.0),
                "forward_ms": 1.0, "reverse_ms": 2.0, "memory_bytes": 0, "n_checkpoints": 1,
                "segment": 1}

    def close(self):
        pass


def test_gradients_follow_input_device_and_dtype():
    amp = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float32, requires_grad=True)
    phase = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    fg0 = torch.zeros((4, 5), dtype=torch.float32, requires_grad=True)
    J = ag.focal_objective(amp, phase, fg0, _StubSim, (1, 2, 1, 2), 10, 5)
    assert J.dtype == torch.float64 and J.device == amp.device and float(J) == 2.5
    (2.0 * J).backward()
    assert amp.grad.dtype == torch.float32 and amp.grad.device == amp.device
    assert phase.grad.dtype == torch.float64
    assert fg0.grad.dtype == torch.float32 and fg0.grad.shape == (4, 5)
    np.testing.assert_allclose(amp.grad.numpy(), [0.0, 6.0, 12.0])
    np.testing.assert_allclose(phase.grad.numpy(), [-2.0, -2.0, -2.0])
    np.testing.assert_allclose(fg0.grad.numpy(), 14.0)


def test_without_void_fraction_map():
    amp = torch.ones(2, dtype=torch.float64, requires_grad=True)
    phase = torch.ones(2, dtype=torch.float64, requires_grad=True)

    class NoFg(_StubSim):
        def adjoint(self, *a, **k):
            self.fg = np.zeros((1, 1))
            return super().adjoint(*a, **k)

    J = ag.focal_objective(amp, phase, None, NoFg, (1, 1, 1, 1), 3, 1)
    J.backward()
    np.testing.assert_allclose(amp.grad.numpy(), [0.0, 3.0])
