"""Helpers for the GPU parity tests: move numpy arrays to the device and compare."""
import numpy as np


def dev_array(a, device="cuda:0"):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0))


def hermite_errors(got, ref, n_points, deriv_scale):
    """Per-plane errors of a Hermite field [f, fx, fy, fxy] (n_points each).

    Plane 0 is measured relative to max|f|.  A derivative plane of order m is measured relative
    to max|f| * deriv_scale**m: for a band-limited field deriv_scale = k_max = pi/h bounds
    max|f'| / max|f| (Bernstein), and rounding in mode k is amplified by k**m; for a map built by
    the footpoint stencil deriv_scale = 1/eps, the amplification of footpoint rounding.
    """
    got = np.asarray(got).ravel()
    ref = np.asarray(ref).ravel()
    f_scale = np.max(np.abs(ref[:n_points]))
    f_scale = f_scale if f_scale > 0 else 1.0
    errs = []
    for p, m in enumerate((0, 1, 1, 2)):
        d = np.max(np.abs(got[p * n_points:(p + 1) * n_points] - ref[p * n_points:(p + 1) * n_points]))
        errs.append(float(d / (f_scale * deriv_scale ** m)))
    return errs
