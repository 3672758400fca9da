"""torch.autograd wrapper of the closest-hit kernel and its adjoint.

    t, beta, gamma, tri = closest_hit(dscene, vertices, origins, directions)

`vertices` (n_vert, 3) must hold the positions the device scene was
committed with (its BVH and triangle records are built from them); they are
an input only so that gradients can flow to them.  The forward pass traces
with the gfx950 kernel; the backward pass runs hiprt_adjoint_closest with the
forward's hit set frozen, recomputing each intersection from its triangle.
Triangle ids are returned as a non-differentiable tensor; rays that miss
have t = FLT_MAX and zero gradients.
"""
import torch

import hiprt

FLT_MAX = 3.4028234663852886e38


def _rays(origins, directions, tmin=1e-4):
    n = origins.shape[0]
    rays = torch.empty((8, n), dtype=torch.float32, device=origins.device)
    rays[0:3] = origins.detach().t()
    rays[3] = tmin
    rays[4:7] = directions.detach().t()
    rays[7] = FLT_MAX
    return rays


class ClosestHit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices, origins, directions, dscene, tracer):
        rays = _rays(origins, directions)
        n = rays.shape[1]
        hits = hiprt.new_hits(n, device=origins.device)
        hiprt.trace_closest(dscene, rays, hits, tracer=tracer)
        ctx.dscene = dscene
        ctx.n_vert = vertices.shape[0]
        ctx.save_for_backward(rays, hits["t"], hits["u"], hits["v"], hits["id"])
        tri = hits["id"].clone()
        ctx.mark_non_differentiable(tri)
        return hits["t"].clone(), hits["u"].clone(), hits["v"].clone(), tri

    @staticmethod
    def backward(ctx, g_t, g_u, g_v, g_tri):
        rays, t, u, v, tri = ctx.saved_tensors
        n = rays.shape[1]
        dev = rays.device
        zeros = lambda: torch.zeros(n, dtype=torch.float32, device=dev)
        g_t = zeros() if g_t is None else g_t.contiguous().float()
        g_u = zeros() if g_u is None else g_u.contiguous().float()
        g_v = zeros() if g_v is None else g_v.contiguous().float()
        gv = torch.zeros((ctx.n_vert, 3), dtype=torch.float32, device=dev)
        go = torch.empty((n, 3), dtype=torch.float32, device=dev)
        gd = torch.empty((n, 3), dtype=torch.float32, device=dev)
        hits = {"t": t, "u": u, "v": v, "id": tri}
        hiprt.adjoint_closest(ctx.dscene, rays, hits, g_t, g_u, g_v, grad_vertices=gv, grad_origin=go, grad_dir=gd)
        return gv, go, gd, None, None


def closest_hit(dscene, vertices, origins, directions, tracer="persistent-while-while"):
    return ClosestHit.apply(vertices, origins, directions, dscene, tracer)
