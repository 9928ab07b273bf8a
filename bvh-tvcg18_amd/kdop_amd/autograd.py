"""torch.autograd harness of the narrow-phase penetration measure.

PenetrationMeasure.apply(pos, quat, ctx, bodies, pairs) refits the bodies to
the poses (pos: (n, 3), quat: (n, 4) as x y z w, float32 on the device; row i
is the pose of bodies[i]), runs the tandem traversal with the narrow phase on
`pairs` (body ids, (m, 2)) and returns Phi per pair: the sum of the depths of
the pair's reported contacts.  backward() runs kdop_narrow_adjoint with the
incoming gradient as the per-pair seed, so dL/dpos and dL/dquat come from the
hand-written FP64 adjoint kernel, with the contact set and every discrete
choice frozen at the forward pass (Phi is piecewise smooth; the adjoint is the
gradient of the smooth piece the forward pass is in).  The quaternion is used
literally (r = Im(q r0 q*) + p, not renormalised).

One context holds one forward pass: refitting or traversing the context again
before backward() raises, since the adjoint reads the context's last
traversal and poses.
"""
import numpy as np
import torch


class PenetrationMeasure(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pos, quat, kctx, bodies, pairs):
        dev = torch.device("cuda", kctx.device)
        p = pos.detach().to(device=dev, dtype=torch.float32).contiguous()
        q = quat.detach().to(device=dev, dtype=torch.float32).contiguous()
        ids = np.ascontiguousarray(np.asarray(bodies, np.uint32).reshape(-1))
        if p.shape != (ids.shape[0], 3) or q.shape != (ids.shape[0], 4):
            raise ValueError(f"pos {tuple(p.shape)} / quat {tuple(q.shape)} do not match {ids.shape[0]} bodies")
        torch.cuda.current_stream(dev).synchronize()  # the pose tensors are written
        kctx.refit_device(ids, p.data_ptr(), q.data_ptr())
        kctx.tandem_traversal(pairs, narrow=True)
        phi, _, _ = kctx.narrow_adjoint(ids)
        fctx.kctx = kctx
        fctx.ids = ids
        fctx.generation = kctx.generation
        fctx.dtypes = (pos.dtype, quat.dtype)
        return phi.clone()

    @staticmethod
    def backward(fctx, gphi):
        kctx = fctx.kctx
        if kctx.generation != fctx.generation:
            raise RuntimeError("the context was refitted or traversed again after this forward pass")
        _, gp, gq = kctx.narrow_adjoint(fctx.ids, dphi=gphi)
        return gp.to(fctx.dtypes[0]), gq.to(fctx.dtypes[1]), None, None, None


def penetration_measure(ctx, bodies, pairs, pos, quat):
    """Phi per pair (differentiable in pos and quat), see PenetrationMeasure."""
    return PenetrationMeasure.apply(pos, quat, ctx, bodies, pairs)
