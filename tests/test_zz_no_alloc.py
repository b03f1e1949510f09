"""cmm_psi_from_vort_noalloc (the psi solve with its workspace in the library's static device array)
against cmm_psi_from_vort on the same inputs, through the C-ABI: two calls in a row, the second on
smaller grids, each equal to the allocating entry point bit for bit (same kernels, same twiddle
tables); a workspace beyond the array is rejected before anything runs; and the calls leave the
device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def inputs(cmm, torch, n_map, n_vort, seed):
    """a smooth perturbed map and a random vorticity Hermite field on the fine (= vort) grid"""
    import oracle_lib as O
    rng = np.random.default_rng(seed)
    cx, cy = O.identity_map(O.grid(n_map))
    n = n_map * n_map
    h = O.TWO_PI / float(np.float32(n_map))
    xs = np.arange(n_map) * h
    X, Y = np.meshgrid(xs, xs)
    a, b = rng.uniform(0.05, 0.2, 2)
    cx[:n] += a * np.sin(Y).ravel()
    cx[2 * n:3 * n] += a * np.cos(Y).ravel()
    cy[:n] += b * np.sin(X).ravel()
    cy[n:2 * n] += b * np.cos(X).ravel()
    w = torch.randn(4 * n_vort * n_vort, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(seed))
    return torch.from_numpy(cx).cuda(), torch.from_numpy(cy).cuda(), w


def solve(cmm, torch, noalloc, gm, gv, gp, cx, cy, w, fcut, out=None):
    L = cmm.lib()
    if out is None:
        out = torch.full((4 * gp.nx * gp.ny,), float("nan"), dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr())
    if noalloc:
        rc = L.cmm_psi_from_vort_noalloc(gm, gv, gv, gp, p(cx), p(cy), p(w), 0, fcut, p(out), st)
    else:
        rc = L.cmm_psi_from_vort(gm, gv, gv, gp, p(cx), p(cy), p(w), 0, fcut, p(out), None, 0, st)
    torch.cuda.synchronize()
    return rc, out


def test_noalloc_psi_matches_and_allocates_nothing(dev):
    import cmm
    import torch
    cases = [(1024, 4096, 2048, 256.0, 1), (128, 512, 256, 40.5, 2)]   # the second call is smaller
    taken = []
    for k, (nm, nv, np_, fcut, seed) in enumerate(cases):
        gm, gv, gp = cmm.grid(nm), cmm.grid(nv), cmm.grid(np_)
        need = cmm.lib().cmm_psi_workspace_bytes(gv, gp, fcut)
        print(f"call {k}: map {nm}^2 vort {nv}^2 psi {np_}^2 cut {fcut}: workspace {need} bytes", flush=True)
        assert need <= 201195520
        cx, cy, w = inputs(cmm, torch, nm, nv, seed)
        rc, ref = solve(cmm, torch, False, gm, gv, gp, cx, cy, w, fcut)
        assert rc == 0, cmm.lib().cmm_last_error()
        got = torch.full_like(ref, float("nan"))
        if k == 0:
            # the first call on the device sets the array up; memory is measured around the later ones
            rc, _ = solve(cmm, torch, True, gm, gv, gp, cx, cy, w, fcut, got)
            assert rc == 0, cmm.lib().cmm_last_error()
            got.fill_(float("nan"))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc, _ = solve(cmm, torch, True, gm, gv, gp, cx, cy, w, fcut, got)
        assert rc == 0, cmm.lib().cmm_last_error()
        free1 = torch.cuda.mem_get_info()[0]
        taken.append(free0 - free1)
        g, r = got.cpu().numpy(), ref.cpu().numpy()
        print(f"call {k}: max |noalloc - allocating| = {np.max(np.abs(g - r)):.3e}; device memory taken by the call:"
              f" {free0 - free1} bytes", flush=True)
        assert np.isfinite(g).all() and np.array_equal(g, r)
        del cx, cy, w, got, ref
    # nothing but the call ran between the two readings: it must not have taken device memory
    assert max(taken) <= (1 << 20), taken


def test_noalloc_psi_rejects_a_workspace_beyond_the_array(dev):
    import cmm
    import torch
    gm, gv, gp = cmm.grid(256), cmm.grid(4096), cmm.grid(2048)
    fcut = 2048.0        # kx_max 1024: spectral buffers of 134 MB plus the 128 MiB sampled field
    assert cmm.lib().cmm_psi_workspace_bytes(gv, gp, fcut) > 201195520
    # full-size buffers, although the call must stop before reading or writing any of them
    cx = torch.zeros(4 * 256 * 256, dtype=torch.float64, device="cuda")
    w = torch.zeros(4 * 4096 * 4096, dtype=torch.float64, device="cuda")
    out = torch.full((4 * 2048 * 2048,), 7.0, dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr())
    rc = cmm.lib().cmm_psi_from_vort_noalloc(gm, gv, gv, gp, p(cx), p(cx), p(w), 0, fcut, p(out), st)
    msg = cmm.lib().cmm_last_error().decode()
    print(rc, msg)
    assert rc != 0 and "exceeds the static scratch array" in msg
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == 7.0).all()
