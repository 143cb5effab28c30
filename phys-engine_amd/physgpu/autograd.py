"""torch.autograd binding of the narrow-phase adjoint (a thin, check-only
layer: every number comes from libphysgpu.so; torch only holds device memory).

    normal, depth, point = contact_geometry(pose, pipe, algorithm)

`pose` is a float32 CUDA tensor [7, n] (px py pz qx qy qz qw per collider) of
the colliders resident in `pipe`.  The forward runs the collision step on
these poses and returns, for the step's manifolds, normal [3, m], depth
[4, m] and world contact points [12, m] (zero beyond each manifold's point
count).  The backward is physgpu_contact_adjoint: gradients of the three
outputs with respect to positions and quaternions (quaternion gradients
tangent to the unit sphere); hull manifolds contribute nothing.

torch ships its own HIP runtime.  Initialise torch's CUDA before the first
physgpu call in the process: libphysgpu.so then binds to torch's runtime and
the two share device memory.  (Loaded the other way round, torch sees no
device, and importing this module raises.)
"""
import torch

from . import BroadPhaseAlgorithm

if not torch.cuda.is_available():
    raise RuntimeError("physgpu.autograd needs torch with a visible HIP device; initialise "
                       "torch's CUDA before loading libphysgpu.so in this process")


class ContactGeometry(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pose, pipe, algorithm):
        if not pose.is_cuda or pose.dtype != torch.float32 or pose.shape != (7, pipe.count):
            raise ValueError(f"pose must be a float32 CUDA tensor [7, {pipe.count}]")
        pose = pose.contiguous()
        torch.cuda.current_stream().synchronize()
        pipe.update_poses_device(pose.data_ptr())
        _, m = pipe.collide(algorithm=algorithm)
        dev = pose.device
        normal = torch.zeros(3, m, device=dev, dtype=torch.float32)
        depth = torch.zeros(4, m, device=dev, dtype=torch.float32)
        point = torch.zeros(12, m, device=dev, dtype=torch.float32)
        if m:
            torch.cuda.current_stream().synchronize()
            pipe.contact_geometry_device(m, normal.data_ptr(), depth.data_ptr(), point.data_ptr())
        ctx.pipe = pipe
        ctx.m = m
        ctx.n = pose.shape[1]
        return normal, depth, point

    @staticmethod
    def backward(ctx, g_normal, g_depth, g_point):
        dev = g_normal.device if g_normal is not None else g_depth.device
        grad = torch.zeros(7, ctx.n, device=dev, dtype=torch.float32)
        if ctx.m:
            def prep(g, rows):
                if g is None:
                    return torch.zeros(rows, ctx.m, device=dev, dtype=torch.float32)
                return g.to(torch.float32).contiguous()
            gn, gd, gp = prep(g_normal, 3), prep(g_depth, 4), prep(g_point, 12)
            torch.cuda.current_stream().synchronize()
            ctx.pipe.contact_adjoint_device(ctx.m, gn.data_ptr(), gd.data_ptr(), gp.data_ptr(),
                                            grad.data_ptr())
        return grad, None, None


def contact_geometry(pose, pipe, algorithm=BroadPhaseAlgorithm.UniformGrid):
    return ContactGeometry.apply(pose, pipe, int(algorithm))
