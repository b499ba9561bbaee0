"""`StirRollout`: the checkpointed rollout as a torch.autograd.Function.

    forces = StirRollout.apply(params, engine)      # params [B, 4] float64, forces [B, T, 3]
    loss = some_torch_function(forces); loss.backward()
    params.grad                                      # dL/dparams from the HIP adjoint

The forward runs ds_reset + ds_forward over the whole trajectory for every instance of the
engine; the backward hands dL/dforces to the library's checkpointed adjoint.  Tensors live on
the host (the force series and the parameters are a few numbers per instance); the particle
state never leaves the GPU.  An engine holds one recorded rollout: a second apply on the same
engine, or new parameters, before loss.backward() makes that backward raise instead of
returning the gradient of another rollout.  This replaces the reference's ti.ad.Tape over forward(s) calls
(src/gradient_descent_module.py:103-107) for torch-side losses and optimisers.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import Engine


class StirRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params: torch.Tensor, engine: Engine) -> torch.Tensor:
        if params.is_cuda:
            # the library's launches follow the work torch queued on its current stream
            engine.set_stream(torch.cuda.current_stream(params.device).cuda_stream)
        p = params.detach().to("cpu", torch.float64).numpy()
        engine.set_params(p.reshape(engine.B, 4))
        engine.reset()
        engine.forward()
        F = engine.forces()
        ctx.engine = engine
        ctx.generation = engine.generation
        return torch.from_numpy(F).to(dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(ctx, grad_forces: torch.Tensor):
        if ctx.engine.generation != ctx.generation:
            raise RuntimeError("StirRollout.backward: the engine was reset or given new parameters after this "
                               "rollout was recorded; its adjoint would belong to another rollout")
        g = grad_forces.detach().to("cpu", torch.float64).numpy()
        d = ctx.engine.backward(np.ascontiguousarray(g))
        return torch.from_numpy(d).to(dtype=torch.float64, device=grad_forces.device), None


def rollout(params: torch.Tensor, engine: Engine) -> torch.Tensor:
    return StirRollout.apply(params, engine)
