"""Thin torch.autograd binding of the batched estimator and its adjoint.

    values = Polyspectra.apply(field, batch)

field is a real N^3 CUDA tensor (float32 or float64) and batch a
polyspectrum_amd.Batch of the context that holds the field's slot.  The
forward uploads the field into the slot, transforms it and evaluates every
configuration (ps_load_field + ps_batch_execute); the backward calls the
hand-written adjoint (ps_batch_adjoint) with the incoming gradient as weights.
torch is only the carrier of device memory here: all arithmetic runs in the
HIP library, on its own stream, which is synchronised with torch's current
stream at both ends.
"""
import numpy as np
import torch


class Polyspectra(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, batch, slot=0):
        if not field.is_cuda:
            raise ValueError("field must be a CUDA tensor")
        f = field.detach().contiguous()
        torch.cuda.current_stream(f.device).synchronize()
        batch._ctx.load_field(f, slot)
        vals, _, _ = batch.execute()
        ctx.batch = batch
        ctx.slot = slot
        ctx.shape = field.shape
        ctx.dtype = field.dtype
        return torch.as_tensor(vals, dtype=torch.float64, device=field.device)

    @staticmethod
    def backward(ctx, grad_values):
        w = grad_values.detach().to("cpu", torch.float64).numpy()
        g = torch.empty(ctx.shape, dtype=ctx.dtype, device=grad_values.device)
        torch.cuda.current_stream(g.device).synchronize()
        ctx.batch.adjoint(np.ascontiguousarray(w), slot=ctx.slot, out=g)
        return g, None, None
