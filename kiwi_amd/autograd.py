"""torch.autograd view of the Cartesian sph-to-grid for fitting and gradient checks.

SphToGrid.apply(m, x, y, z, h, vx, vy, vz, ctx, domain, res, precision) returns the per-cell sums
(mass, px, py, pz) of kiwi_hip_sph_to_grid; its backward calls kiwi_hip_sph_to_grid_adjoint, which
recomputes the kernel integrals (nothing per pair is stored between the two passes).  All tensors are
float64 on the device.  PyTorch is plumbing here (device memory and the tape); the computation is the
HIP library's.  The library's streams do not wait for torch's: kiwi_amd.hip synchronises torch's current
stream before every call with device tensors (the zero-filled outputs, inputs an optimiser step has just
written, the upstream gradients), and the calls are synchronous.
"""
import torch

from . import hip


class SphToGrid(torch.autograd.Function):
    @staticmethod
    def forward(fctx, m, x, y, z, h, vx, vy, vz, kctx, domain, res, precision=hip.FP64):
        cols = {"x": x, "y": y, "z": z, "vx": vx, "vy": vy, "vz": vz, "m": m, "h": h}
        parts = {k: v.detach().contiguous() for k, v in cols.items()}
        nc = int(res[0]) * int(res[1]) * int(res[2])
        out = tuple(torch.zeros(nc, dtype=torch.float64, device=x.device) for _ in range(4))
        kctx.sph_to_grid(parts, domain, res, precision=precision, out=out)
        torch.cuda.synchronize(x.device)
        fctx.kctx, fctx.domain, fctx.res, fctx.precision = kctx, domain, res, precision
        fctx.save_for_backward(m, x, y, z, h, vx, vy, vz)
        return out

    @staticmethod
    def backward(fctx, g_mass, g_px, g_py, g_pz):
        m, x, y, z, h, vx, vy, vz = fctx.saved_tensors
        parts = {"x": x.detach().contiguous(), "y": y.detach().contiguous(), "z": z.detach().contiguous(),
                 "vx": vx.detach().contiguous(), "vy": vy.detach().contiguous(), "vz": vz.detach().contiguous(),
                 "m": m.detach().contiguous(), "h": h.detach().contiguous()}
        nc = g_mass.numel() if g_mass is not None else int(fctx.res[0]) * int(fctx.res[1]) * int(fctx.res[2])
        zeros = torch.zeros(nc, dtype=torch.float64, device=x.device)
        gin = tuple((g.contiguous() if g is not None else zeros) for g in (g_mass, g_px, g_py, g_pz))
        out = {f: torch.zeros_like(x) for f in hip.GRAD_FIELDS}
        fctx.kctx.sph_to_grid_adjoint(parts, fctx.domain, fctx.res, gin, precision=fctx.precision, out=out)
        torch.cuda.synchronize(x.device)
        return (out["m"], out["x"], out["y"], out["z"], out["h"], out["vx"], out["vy"], out["vz"],
                None, None, None, None)


def sph_to_grid(m, x, y, z, h, vx, vy, vz, kctx, domain, res, precision=hip.FP64):
    return SphToGrid.apply(m, x, y, z, h, vx, vy, vz, kctx, domain, res, precision)
