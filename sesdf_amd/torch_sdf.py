"""PyTorch-ROCm front-end: SDF sampling of the SES as a differentiable function of the atoms.

``SurfaceSDF.apply(atoms, points, surface, band, store, exterior)`` builds the complete surface of
``atoms`` ([N, 4] float32 on the GPU: x, y, z, vdW radius, input order) with the handle ``surface``
(a ``sesdf_amd.Sesdf`` created with the same atom count), and returns SDF(points) ([Q] float32).
Its backward gives d(sum_q g_q SDF(q_q))/d(atoms) with the patch topology of the forward build held
fixed (the hand-written HIP backward kernel behind the C ABI). ``store`` trades memory for time:
  True       keeps the winning element of every query from the forward pass (8 bytes per query);
  False      keeps the topology and re-derives the winners in the backward pass;
  "release"  frees the whole build (neighbour lists, circles, patches, segments: the device
             buffers of the handle) after the forward pass; the backward pass builds the same
             atoms again (deterministic: the same topology) and re-derives the winners.

The handle may be rebuilt between forward and backward (two forwards on one handle before one
backward, several conformers in one loss): the backward then builds the forward's atoms again
first. Builds are deterministic, so the stored winners index the same elements.
"""
import torch

from .surface import Sesdf


class SurfaceSDF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, atoms, points, surface, band=2.0, store=True, exterior=False):
        if not (atoms.is_cuda and points.is_cuda):
            raise ValueError("atoms and points must be on the GPU")
        atoms = atoms.detach().contiguous().float()
        pts = points.detach().float()
        if pts.shape[-1] == 3:
            pts = torch.nn.functional.pad(pts, (0, 1))
        pts = pts.contiguous()
        if atoms.shape[0] != surface.atom_nb:
            raise ValueError("surface handle was created for a different atom count")
        if store not in (True, False, "release"):
            raise ValueError('store must be True, False or "release"')
        keep = store is True
        torch.cuda.current_stream().synchronize()
        surface.build_device(atoms.data_ptr())
        q = pts.shape[0]
        out = torch.empty(q, device=pts.device, dtype=torch.float32)
        codes = torch.empty(q, device=pts.device, dtype=torch.int64) if keep else None
        surface.sample_points_ptr(pts.data_ptr(), q, out.data_ptr(), band=band, exterior=exterior,
                                  d_codes=codes.data_ptr() if keep else None)
        _sync(surface)
        ctx.surface, ctx.band, ctx.exterior, ctx.codes = surface, band, exterior, codes
        ctx.serial = surface.counts()["build_serial"]
        ctx.released = store == "release"
        if ctx.released:
            surface.release()
        ctx.save_for_backward(pts, atoms)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        pts, atoms = ctx.saved_tensors
        surface = ctx.surface
        if ctx.released or surface.counts()["build_serial"] != ctx.serial:
            # the handle holds another build now: build the forward's atoms again
            torch.cuda.current_stream().synchronize()
            surface.build_device(atoms.data_ptr())
            _sync(surface)
        w = grad_out.detach().contiguous().float()
        n = surface.atom_nb
        grad = torch.zeros(n, 4, device=pts.device, dtype=torch.float32)
        scratch = torch.empty(n, 4, device=pts.device, dtype=torch.float32)
        torch.cuda.current_stream().synchronize()
        surface.sdf_backward_ptr(pts.data_ptr(), w.data_ptr(), pts.shape[0], grad.data_ptr(), scratch.data_ptr(),
                                 band=ctx.band, exterior=ctx.exterior,
                                 d_codes=ctx.codes.data_ptr() if ctx.codes is not None else None)
        _sync(surface)
        return grad, None, None, None, None, None


def _sync(surface):
    import sesdf_amd._lib as L

    L.device_synchronize()


def sdf(atoms, points, surface=None, band=2.0, store=True, exterior=False, probe_radius=1.4, **kw):
    """Convenience wrapper: creates a handle when none is given."""
    if surface is None:
        surface = Sesdf(atoms.detach().cpu().numpy(), probe_radius=probe_radius, build_surface=False, **kw)
    return SurfaceSDF.apply(atoms, points, surface, band, store, exterior)
