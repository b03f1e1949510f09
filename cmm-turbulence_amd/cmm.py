"""Python binding of libcmm.so (the MI355X CMM engine) over its C-ABI (include/cmm.h).

Device buffers are torch CUDA tensors of dtype float64 (torch is only the allocator and the
stream); every call goes straight to the HIP library.  There is no CPU fallback: loading fails
loudly when the library has not been built.

The operator names follow the reference's host interface (Cuda-CMM/src/simulation/cmm-simulation-host.h:23-50,
Cuda-CMM/src/numerical/cmm-particles.h:29-32, Cuda-CMM/src/numerical/cmm-fft.h:51).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libcmm.so")
TWO_PI = 6.283185307179586476925286766559005768394338798750211641949889184615632812572

# initial-condition ids (include/cmm_types.h); names as in Cuda-CMM/src/ui/settings.h:313-324
IC = {"4_nodes": 0, "quadropole": 1, "one_vortex": 2, "two_votices": 3, "three_vortices": 4,
      "single_shear_layer": 5, "tanh_shear_layer": 6, "shielded_vortex": 9, "two_cosine": 10,
      "vortex_sheets": 11, "double_shear_layer": 100, "random_spectrum": 101,
      # distribution functions of the Vlasov-Poisson flow, Cuda-CMM/src/ui/settings.h:300-303
      "landau_damping": 200, "two_stream": 201, "nonlin_landau_damping": 202,
      # ensemble conditions with amplitude ic_params[0] (no reference counterpart)
      "landau_amp": 203, "two_stream_v2": 204}
SIM = {"cmm_euler_2d": 0, "cmm_vlasov_poisson_1d": 1}
# time integration names, Cuda-CMM/src/ui/settings.h:386-398
TIME_INTEGRATION = {"EulerExp": 10, "RK1": 10, "AB2": 20, "RK2": 21, "Heun": 21, "RK3": 30, "RK3Mod": 31,
                    "RK4": 40, "RK4Mod": 41}
MAP_UPDATE_ORDER = {"2nd": 0, "4th": 1, "6th": 2}


class CmmError(RuntimeError):
    pass


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double)]

    @property
    def n(self):
        return self.nx * self.ny


class AdvectCoeffs(C.Structure):
    _fields_ = [("L1", C.c_double * 4), ("L12", C.c_double * 4), ("c1", C.c_double * 12),
                ("cx", C.c_double * 12), ("cy", C.c_double * 12), ("cxy", C.c_double * 12),
                ("dt", C.c_double), ("eps", C.c_double), ("time_integration_num", C.c_int),
                ("map_update_order_num", C.c_int), ("l_order", C.c_int), ("direction", C.c_int)]


class EulerParams(C.Structure):
    _fields_ = [("g_coarse", Grid), ("g_fine", Grid), ("g_psi", Grid), ("g_vort", Grid),
                ("dt", C.c_double),
                ("time_integration_num", C.c_int), ("map_update_order_num", C.c_int), ("l_order", C.c_int),
                ("lagrange_init_higher_order", C.c_int),
                ("map_epsilon", C.c_double), ("incomp_threshold", C.c_double), ("freq_cut_psi", C.c_double),
                ("molly_stencil", C.c_int), ("skip_remapping", C.c_int),
                ("ic", C.c_int), ("ic_params", C.c_double * 10),
                ("discrete", C.c_int), ("g_disc", Grid), ("disc_real", C.POINTER(C.c_double)),
                ("seed", C.c_uint64), ("stack_capacity", C.c_int), ("sim", C.c_int), ("forward_map", C.c_int),
                ("precision", C.c_int)]


class StepInfo(C.Structure):
    _fields_ = [("t", C.c_double), ("incomp", C.c_double), ("remapped", C.c_int), ("stack_size", C.c_int),
                ("saturated", C.c_int), ("incomp_forward", C.c_double), ("invertibility", C.c_double)]


PROF_NAMES = ["advect", "incomp", "psi_row_r2c", "psi_col", "psi_row_c2r", "particles", "remap", "psi_combine"]
PROF_SLOTS = len(PROF_NAMES)
FIELD = {"chix": 0, "chiy": 1, "psi": 2, "w_fine": 3, "w_coarse": 4, "disc": 5, "chix_f": 6, "chiy_f": 7}

# every symbol include/cmm.h declares
SYMBOLS = ["cmm_version", "cmm_last_error", "cmm_device_info", "cmm_lagrange_coeff", "cmm_advect_coeffs_fill",
           "cmm_map_init_identity", "cmm_advect_workspace_bytes", "cmm_map_advect", "cmm_incomp_error", "cmm_invert_error", "cmm_compose_positions",
           "cmm_stack_pack", "cmm_compose_positions_packed", "cmm_forward_positions", "cmm_forward_points",
           "cmm_sample_init", "cmm_translate_workspace_bytes", "cmm_translate_ic", "cmm_hermite_workspace_bytes",
           "cmm_hermite_from_real", "cmm_hermite_from_spectrum", "cmm_psi_workspace_bytes", "cmm_psi_from_vort", "cmm_psi_from_vort_noalloc",
           "cmm_spectral_workspace_bytes", "cmm_spectral_from_real", "cmm_sample_vort", "cmm_vp_psi_workspace_bytes", "cmm_vp_psi_from_f", "cmm_particles_advect", "cmm_euler_create", "cmm_euler_destroy", "cmm_euler_stream",
           "cmm_euler_add_particles", "cmm_euler_add_particle_set", "cmm_euler_get_particle_velocities", "cmm_euler_init", "cmm_euler_step", "cmm_euler_get_field",
           "cmm_euler_get_particles", "cmm_euler_diagnostics", "cmm_euler_vp_diagnostics", "cmm_euler_sample", "cmm_euler_vp_sample", "cmm_euler_zoom", "cmm_euler_restart", "cmm_euler_get_stack_map",
           "cmm_euler_restart_forward", "cmm_euler_get_forward_stack_map", "cmm_euler_request_remap",
           "cmm_euler_forward_positions", "cmm_euler_forward_points", "cmm_euler_state", "cmm_euler_profile", "cmm_euler_profile_read",
           "cmm_map_to_split", "cmm_map_from_split", "cmm_hermite_to_split", "cmm_map_advect_split_workspace_bytes",
           "cmm_map_advect_split", "cmm_incomp_error_split", "cmm_psi_from_vort_split",
           "cmm_batch_create", "cmm_batch_destroy", "cmm_batch_step", "cmm_batch_stats", "cmm_batch_set_grouping", "cmm_batch_grouped_steps",
           "cmm_batch_autotune", "cmm_batch_set_mode", "cmm_euler_autotune", "cmm_euler_set_check_tile",
           "cmm_euler_set_check_fusion",
           "cmm_euler_get_check_fusion", "cmm_euler_set_psi_precombine",
           "cmm_euler_get_psi_precombine",
           "cmm_cnn_forward", "cmm_cnn_backward_data", "cmm_cnn_backward_data_relu", "cmm_cnn_backward_weights",
           "cmm_cnn_relu_mask",
           "cmm_cnn_fill", "cmm_cnn_weights_bf16", "cmm_cnn_adamw", "cmm_cnn_input", "cmm_cnn_target", "cmm_cnn_mse",
           "cmm_cnn_correction", "cmm_extrapolate", "cmm_dev_norms", "cmm_euler_copy_psi_level",
           "cmm_euler_psi_level", "cmm_euler_set_psi_correction"]
# engine precision forms (CMM_PRECISION_*)
PRECISION = {"fp64": 0, "split": 1}

_lib = None
DP = C.POINTER(C.c_double)
VP = C.c_void_p


def lib():
    """Load libcmm.so; raises when it is missing (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CmmError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C cmm-turbulence_amd")
        L = C.CDLL(LIB_PATH)
        L.cmm_last_error.restype = C.c_char_p
        L.cmm_lagrange_coeff.restype = C.c_double
        L.cmm_lagrange_coeff.argtypes = [DP, C.c_double, C.c_int, C.c_int, C.c_int]
        L.cmm_advect_coeffs_fill.argtypes = [DP, DP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                             C.POINTER(AdvectCoeffs)]
        L.cmm_map_init_identity.argtypes = [Grid, VP, VP, VP]
        L.cmm_advect_workspace_bytes.restype = C.c_size_t
        L.cmm_advect_workspace_bytes.argtypes = [Grid, Grid, C.POINTER(AdvectCoeffs)]
        L.cmm_map_advect.argtypes = [Grid, Grid, VP, VP, C.POINTER(VP), C.POINTER(AdvectCoeffs), VP, VP, VP,
                                     C.c_size_t, VP]
        L.cmm_incomp_error.argtypes = [Grid, Grid, VP, VP, DP, VP]
        L.cmm_invert_error.argtypes = [Grid, Grid, Grid, VP, VP, VP, VP, DP, VP]
        L.cmm_compose_positions.argtypes = [Grid, Grid, VP, VP, VP, C.c_int, VP, VP]
        L.cmm_forward_positions.argtypes = [Grid, Grid, VP, VP, VP, C.c_int, VP, VP]
        L.cmm_stack_pack.argtypes = [Grid, VP, C.c_int, C.c_int, VP, VP]
        L.cmm_compose_positions_packed.argtypes = [Grid, Grid, VP, VP, VP, C.c_int, C.c_int, VP, VP]
        L.cmm_forward_points.argtypes = [Grid, Grid, VP, VP, VP, C.c_int, C.c_int, C.c_int, C.c_int, VP, VP, VP]
        L.cmm_sample_init.argtypes = [Grid, Grid, VP, VP, VP, C.c_int, C.c_int, DP, C.c_int, Grid, VP, VP, VP]
        L.cmm_translate_workspace_bytes.restype = C.c_size_t
        L.cmm_translate_workspace_bytes.argtypes = [Grid]
        L.cmm_translate_ic.argtypes = [Grid, Grid, VP, VP, VP, C.c_int, C.c_int, DP, C.c_int, Grid, VP, VP, VP,
                                       C.c_size_t, VP]
        L.cmm_hermite_workspace_bytes.restype = C.c_size_t
        L.cmm_hermite_workspace_bytes.argtypes = [Grid]
        L.cmm_hermite_from_real.argtypes = [Grid, VP, VP, VP, C.c_size_t, VP]
        L.cmm_hermite_from_spectrum.argtypes = [Grid, VP, VP, VP, C.c_size_t, VP]
        L.cmm_psi_workspace_bytes.restype = C.c_size_t
        L.cmm_psi_workspace_bytes.argtypes = [Grid, Grid, C.c_double]
        L.cmm_psi_from_vort.argtypes = [Grid, Grid, Grid, Grid, VP, VP, VP, C.c_int, C.c_double, VP, VP,
                                        C.c_size_t, VP]
        L.cmm_psi_from_vort_noalloc.argtypes = [Grid, Grid, Grid, Grid, VP, VP, VP, C.c_int, C.c_double, VP, VP]
        L.cmm_sample_vort.argtypes = [Grid, Grid, Grid, VP, VP, VP, C.c_int, VP, VP]
        L.cmm_spectral_workspace_bytes.restype = C.c_size_t
        L.cmm_spectral_workspace_bytes.argtypes = [Grid]
        L.cmm_spectral_from_real.argtypes = [Grid, VP, C.c_int, VP, VP, C.c_size_t, VP]
        L.cmm_vp_psi_workspace_bytes.restype = C.c_size_t
        L.cmm_vp_psi_workspace_bytes.argtypes = [Grid, Grid]
        L.cmm_vp_psi_from_f.argtypes = [Grid, Grid, Grid, Grid, VP, VP, VP, C.c_int, VP, VP, C.c_size_t, VP]
        L.cmm_particles_advect.argtypes = [Grid, C.POINTER(VP), C.c_int, DP, C.c_double, C.c_int, C.c_int, VP, VP]
        L.cmm_euler_create.argtypes = [C.POINTER(EulerParams), C.POINTER(VP)]
        L.cmm_euler_destroy.argtypes = [VP]
        L.cmm_euler_stream.restype = VP
        L.cmm_euler_stream.argtypes = [VP]
        L.cmm_euler_add_particles.argtypes = [VP, C.c_int, C.c_int, DP, C.POINTER(C.c_int)]
        L.cmm_euler_add_particle_set.argtypes = [VP, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, DP, DP,
                                                 C.POINTER(C.c_int)]
        L.cmm_euler_get_particle_velocities.argtypes = [VP, C.c_int, DP]
        L.cmm_euler_init.argtypes = [VP, DP, C.POINTER(C.c_int)]
        L.cmm_euler_step.argtypes = [VP, C.c_double, C.POINTER(StepInfo)]
        L.cmm_euler_get_field.argtypes = [VP, C.c_int, DP]
        L.cmm_euler_get_particles.argtypes = [VP, C.c_int, DP]
        L.cmm_euler_diagnostics.argtypes = [VP, DP]
        L.cmm_euler_vp_diagnostics.argtypes = [VP, DP]
        L.cmm_euler_sample.argtypes = [VP, Grid, C.c_int, DP]
        L.cmm_euler_vp_sample.argtypes = [VP, Grid, DP]
        L.cmm_euler_zoom.argtypes = [VP, Grid, C.POINTER(C.c_int), C.c_int, C.c_int, DP]
        L.cmm_euler_restart.argtypes = [VP, C.c_double, C.c_int, DP, DP, DP, DP]
        L.cmm_euler_get_stack_map.argtypes = [VP, C.c_int, DP, DP]
        L.cmm_euler_restart_forward.argtypes = [VP, C.c_int, DP, DP, DP]
        L.cmm_euler_get_forward_stack_map.argtypes = [VP, C.c_int, DP, DP]
        L.cmm_euler_request_remap.argtypes = [VP]
        L.cmm_euler_forward_positions.argtypes = [VP, Grid, DP]
        L.cmm_euler_forward_points.argtypes = [VP, Grid, C.c_int, DP, C.c_int, C.c_int, DP]
        L.cmm_euler_state.argtypes = [VP, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.cmm_euler_profile.argtypes = [VP, C.c_int]
        L.cmm_euler_profile_read.argtypes = [VP, DP, C.POINTER(C.c_int)]
        L.cmm_map_to_split.argtypes = [Grid, VP, VP, VP, VP]
        L.cmm_map_from_split.argtypes = [Grid, VP, VP, VP, VP]
        L.cmm_hermite_to_split.argtypes = [Grid, VP, VP, VP]
        L.cmm_map_advect_split_workspace_bytes.restype = C.c_size_t
        L.cmm_map_advect_split_workspace_bytes.argtypes = [Grid, C.POINTER(AdvectCoeffs)]
        L.cmm_map_advect_split.argtypes = [Grid, Grid, VP, C.POINTER(VP), C.POINTER(AdvectCoeffs), VP, VP, C.c_size_t, VP]
        L.cmm_incomp_error_split.argtypes = [Grid, Grid, VP, DP, VP]
        L.cmm_psi_from_vort_split.argtypes = [Grid, Grid, Grid, Grid, VP, VP, C.c_int, C.c_double, VP, VP, C.c_size_t,
                                              VP]
        L.cmm_batch_create.argtypes = [C.POINTER(VP), C.c_int, C.c_int, C.POINTER(VP)]
        L.cmm_batch_destroy.argtypes = [VP]
        L.cmm_batch_step.argtypes = [VP, C.c_double, C.POINTER(StepInfo)]
        L.cmm_batch_stats.argtypes = [VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cmm_batch_set_grouping.argtypes = [VP, C.c_int]
        L.cmm_batch_grouped_steps.argtypes = [VP, C.POINTER(C.c_int)]
        L.cmm_batch_autotune.argtypes = [VP, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.cmm_batch_set_mode.argtypes = [VP, C.c_int]
        L.cmm_euler_autotune.argtypes = [VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.cmm_euler_set_check_tile.argtypes = [VP, C.c_int, C.c_int]
        L.cmm_euler_set_check_fusion.argtypes = [VP, C.c_int]
        L.cmm_euler_get_check_fusion.argtypes = [VP, C.POINTER(C.c_int)]
        L.cmm_euler_set_psi_precombine.argtypes = [VP, C.c_int]
        L.cmm_euler_get_psi_precombine.argtypes = [VP, C.POINTER(C.c_int)]
        L.cmm_cnn_forward.argtypes = [C.c_int, C.c_int, C.c_int, VP, VP, VP, C.c_int, VP, VP]
        L.cmm_cnn_backward_data.argtypes = [C.c_int, C.c_int, C.c_int, VP, VP, VP, VP]
        L.cmm_cnn_backward_data_relu.argtypes = [C.c_int, C.c_int, C.c_int, VP, VP, VP, VP, VP]
        L.cmm_cnn_backward_weights.argtypes = [C.c_int, C.c_int, C.c_int, VP, VP, VP, VP, VP]
        L.cmm_cnn_relu_mask.argtypes = [C.c_long, VP, VP, VP]
        L.cmm_cnn_fill.argtypes = [C.c_long, VP, C.c_float, VP]
        L.cmm_cnn_weights_bf16.argtypes = [C.c_int, VP, VP, VP, VP]
        L.cmm_cnn_adamw.argtypes = [C.c_long, VP, VP, VP, VP, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int, VP]
        L.cmm_cnn_input.argtypes = [C.c_long, VP, VP, VP, VP, VP, VP]
        L.cmm_cnn_target.argtypes = [C.c_long, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]
        L.cmm_cnn_mse.argtypes = [C.c_long, VP, VP, VP, VP, VP]
        L.cmm_cnn_correction.argtypes = [C.c_long, VP, VP, VP, VP, VP]
        L.cmm_extrapolate.argtypes = [C.c_long, VP, VP, VP, VP, VP, VP, VP]
        L.cmm_dev_norms.argtypes = [C.c_long, VP, VP, VP, VP]
        L.cmm_euler_copy_psi_level.argtypes = [VP, C.c_int, C.c_int, VP]
        L.cmm_euler_psi_level.argtypes = [VP, C.c_int, C.POINTER(VP)]
        L.cmm_euler_set_psi_correction.argtypes = [VP, VP, VP]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise CmmError(f"{what} failed ({rc}): {lib().cmm_last_error().decode()}")


def grid(n, ny=None, bounds=(0.0, TWO_PI, 0.0, TWO_PI)):
    return Grid(n, n if ny is None else ny, bounds[0], bounds[1], bounds[2], bounds[3])


def _p(t):
    """device pointer of a torch tensor (None -> NULL)"""
    if t is None:
        return None
    import torch
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous(), "expected contiguous float64 cuda tensor"
    return C.c_void_p(t.data_ptr())


def _pf(t):
    """device pointer of a float32 torch tensor (split-precision records)"""
    import torch
    assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), "expected contiguous float32 cuda tensor"
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _levels(levels):
    arr = (VP * 4)()
    for i, t in enumerate(levels):
        arr[i] = t.data_ptr()
    return arr


def advect_coeffs(t, dt, loop_ctr, l_order, tin, muo, eps, direction=-1):
    import numpy as np
    c = AdvectCoeffs()
    t = np.ascontiguousarray(t, dtype=np.float64)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    _check(lib().cmm_advect_coeffs_fill(t.ctypes.data_as(DP), dt.ctypes.data_as(DP), loop_ctr, l_order, tin, muo,
                                        eps, direction, C.byref(c)), "cmm_advect_coeffs_fill")
    return c


def map_init_identity(g, chix, chiy):
    _check(lib().cmm_map_init_identity(g, _p(chix), _p(chiy), _stream()), "cmm_map_init_identity")


def advect_workspace_bytes(g_map, g_psi, coeffs):
    return lib().cmm_advect_workspace_bytes(g_map, g_psi, C.byref(coeffs))


def advect_using_stream_hermite(g_map, g_psi, chix, chiy, psi_levels, coeffs, chix_new, chiy_new, work=None):
    """work: None interpolates every psi level per evaluation (the reference's order); a device
    buffer of advect_workspace_bytes() bytes lets the call combine the levels first."""
    wb = 0 if work is None else work.numel() * work.element_size()
    _check(lib().cmm_map_advect(g_map, g_psi, _p(chix), _p(chiy), _levels(psi_levels), C.byref(coeffs),
                                _p(chix_new), _p(chiy_new), _p(work) if work is not None else None, wb, _stream()),
           "cmm_map_advect")


def incompressibility_check(g_check, g_map, chix, chiy):
    out = C.c_double()
    _check(lib().cmm_incomp_error(g_map, g_check, _p(chix), _p(chiy), C.byref(out), _stream()), "cmm_incomp_error")
    return out.value


def invertibility_check(g_check, g_back, chix_b, chiy_b, g_fwd, chix_f, chiy_f):
    """max |x - X_b(X_f(x))| over the cell centres of g_check (cmm_invert_error)"""
    out = C.c_double()
    _check(lib().cmm_invert_error(g_back, g_fwd, g_check, _p(chix_b), _p(chiy_b), _p(chix_f), _p(chiy_f),
                                  C.byref(out), _stream()), "cmm_invert_error")
    return out.value


def apply_map_stack(g_target, g_map, chix, chiy, stack, n_stack, xy):
    _check(lib().cmm_compose_positions(g_map, g_target, _p(chix), _p(chiy), _p(stack), n_stack, _p(xy), _stream()),
           "cmm_compose_positions")


# stack record formats (CMM_STACK_RECORDS_*): 64 bytes per node FP64, 48 bytes with FP32 derivatives
STACK_RECORDS = {"fp64": 1, "split": 2}
STACK_RECORD_BYTES = {"fp64": 64, "split": 48}


def stack_pack(g_map, stack, n_maps, fmt, out):
    """plane-major stacked maps into node records (out holds STACK_RECORD_BYTES[fmt] * Nc * n_maps
    bytes)"""
    _check(lib().cmm_stack_pack(g_map, _p(stack), n_maps, STACK_RECORDS[fmt], _p(out), _stream()), "cmm_stack_pack")


def apply_map_stack_packed(g_target, g_map, chix, chiy, packed, n_stack, fmt, xy):
    """apply_map_stack through a stack packed by stack_pack"""
    _check(lib().cmm_compose_positions_packed(g_map, g_target, _p(chix), _p(chiy), _p(packed), n_stack,
                                              STACK_RECORDS[fmt], _p(xy), _stream()), "cmm_compose_positions_packed")


def apply_map_stack_forward(g_target, g_map, chix, chiy, stack, n_stack, xy):
    """forward positions: the oldest stacked forward map first, the current one last"""
    _check(lib().cmm_forward_positions(g_map, g_target, _p(chix), _p(chiy), _p(stack), n_stack, _p(xy), _stream()),
           "cmm_forward_positions")


def apply_map_stack_points_forward(g_bounds, g_map, chix, chiy, stack, n_stack, first_map, apply_current, pos_in,
                                   pos_out):
    """forwarded particles through stacked forward maps first_map.. and then the current one"""
    _check(lib().cmm_forward_points(g_map, g_bounds, _p(chix), _p(chiy), _p(stack), n_stack, first_map,
                                    int(apply_current), pos_in.numel() // 2, _p(pos_in), _p(pos_out), _stream()),
           "cmm_forward_points")


def sample_init(g_map, g_target, chix, chiy, stack, n_stack, ic, out, ic_params=None, disc=None, g_disc=None):
    prm = (C.c_double * 10)(*([0.0] * 10 if ic_params is None else list(ic_params) + [0.0] * (10 - len(ic_params))))
    gd = g_disc if g_disc is not None else g_target
    _check(lib().cmm_sample_init(g_map, g_target, _p(chix), _p(chiy), _p(stack), n_stack, ic,
                                 C.cast(prm, DP), 1 if disc is not None else 0, gd, _p(disc), _p(out), _stream()),
           "cmm_sample_init")


def translate_initial_condition_through_map_stack(g_map, g_fine, chix, chiy, stack, n_stack, ic, w_herm,
                                                  ic_params=None, disc=None, g_disc=None):
    prm = (C.c_double * 10)(*([0.0] * 10 if ic_params is None else list(ic_params) + [0.0] * (10 - len(ic_params))))
    gd = g_disc if g_disc is not None else g_fine
    _check(lib().cmm_translate_ic(g_map, g_fine, _p(chix), _p(chiy), _p(stack), n_stack, ic, C.cast(prm, DP),
                                  1 if disc is not None else 0, gd, _p(disc), _p(w_herm), None, 0, _stream()),
           "cmm_translate_ic")


def hermite_from_real(g, real, herm):
    _check(lib().cmm_hermite_from_real(g, _p(real), _p(herm), None, 0, _stream()), "cmm_hermite_from_real")


def spectral_from_real(g, real, op, herm):
    """Hermite planes of F (op 0), of i_laplacian(F) (op 1) or of laplacian(F) (op 2)."""
    _check(lib().cmm_spectral_from_real(g, _p(real), op, _p(herm), None, 0, _stream()), "cmm_spectral_from_real")


def hermite_from_spectrum(g, spec_cols, herm):
    _check(lib().cmm_hermite_from_spectrum(g, _p(spec_cols), _p(herm), None, 0, _stream()),
           "cmm_hermite_from_spectrum")


def evaluate_stream_hermite(g_map, g_vort, g_fine, g_psi, chix, chiy, w_herm, psi_out, molly_stencil=0,
                            freq_cut_psi=128.0):
    _check(lib().cmm_psi_from_vort(g_map, g_vort, g_fine, g_psi, _p(chix), _p(chiy), _p(w_herm), molly_stencil,
                                   freq_cut_psi, _p(psi_out), None, 0, _stream()), "cmm_psi_from_vort")


def evaluate_potential_from_density_hermite(g_map, g_vort, g_fine, g_psi, chix, chiy, f_herm, psi_out,
                                             molly_stencil=0):
    _check(lib().cmm_vp_psi_from_f(g_map, g_vort, g_fine, g_psi, _p(chix), _p(chiy), _p(f_herm), molly_stencil,
                                   _p(psi_out), None, 0, _stream()), "cmm_vp_psi_from_f")


# ---- split-precision form (no counterpart in the reference): map records of 8 floats per node
# (X - x, dX/dx - 1, dX/dy, d2X/dxdy, Y - y, dY/dx, dY/dy - 1, d2Y/dxdy), FP32 Hermite node records

def map_to_split(g, chix, chiy, m_split):
    _check(lib().cmm_map_to_split(g, _p(chix), _p(chiy), _pf(m_split), _stream()), "cmm_map_to_split")


def map_from_split(g, m_split, chix, chiy):
    _check(lib().cmm_map_from_split(g, _pf(m_split), _p(chix), _p(chiy), _stream()), "cmm_map_from_split")


def hermite_to_split(g, herm, h_split):
    _check(lib().cmm_hermite_to_split(g, _p(herm), _pf(h_split), _stream()), "cmm_hermite_to_split")


def advect_split(g_map, g_psi, m_in, psi_levels, coeffs, m_out):
    import torch
    wb = lib().cmm_map_advect_split_workspace_bytes(g_psi, C.byref(coeffs))
    work = torch.empty(max(wb, 8), dtype=torch.uint8, device=m_in.device)
    _check(lib().cmm_map_advect_split(g_map, g_psi, _pf(m_in), _levels(psi_levels), C.byref(coeffs), _pf(m_out),
                                      C.c_void_p(work.data_ptr()), wb, _stream()), "cmm_map_advect_split")


def incompressibility_check_split(g_check, g_map, m_split):
    out = C.c_double()
    _check(lib().cmm_incomp_error_split(g_map, g_check, _pf(m_split), C.byref(out), _stream()),
           "cmm_incomp_error_split")
    return out.value


def evaluate_stream_hermite_split(g_map, g_vort, g_fine, g_psi, m_split, w_split, psi_out, molly_stencil=0,
                                  freq_cut_psi=128.0):
    _check(lib().cmm_psi_from_vort_split(g_map, g_vort, g_fine, g_psi, _pf(m_split), _pf(w_split), molly_stencil,
                                         freq_cut_psi, _p(psi_out), None, 0, _stream()), "cmm_psi_from_vort_split")


def vp_grid(n, nv=None):
    """(x, v) grid of the Vlasov-Poisson flow on [0, 4 pi) x [-2.5 pi, 2.5 pi), the reference's
    bounds {0, 4, -2.5, 2.5} times pi (Cuda-CMM/src/simulation/cmm-vlasov1d.cu:64-75)."""
    return grid(n, nv, (0.0, 2 * TWO_PI, -1.25 * TWO_PI, 1.25 * TWO_PI))


def sample_vort(g_map, g_target, g_fine, chix, chiy, w_herm, out, molly_stencil=0):
    _check(lib().cmm_sample_vort(g_map, g_target, g_fine, _p(chix), _p(chiy), _p(w_herm), molly_stencil, _p(out),
                                 _stream()), "cmm_sample_vort")


def particles_advect(g_psi, psi_levels, l_order, Lp12, dt, scheme, pos):
    Lp = (C.c_double * 4)(*(list(Lp12) + [0.0] * (4 - len(Lp12))))
    _check(lib().cmm_particles_advect(g_psi, _levels(psi_levels), l_order, C.cast(Lp, DP), dt, scheme,
                                      pos.numel() // 2, _p(pos), _stream()), "cmm_particles_advect")


class Euler:
    """The device Euler time loop (cmm_euler_* in include/cmm.h)."""

    def __init__(self, g_coarse, g_fine, g_psi, g_vort, dt, time_integration="RK3", map_update_order="4th",
                 l_order=None, lagrange_init_higher_order=True, map_epsilon=1e-3, incomp_threshold=1e-4,
                 freq_cut_psi=128.0, molly_stencil=0, skip_remapping=False, initial_condition="4_nodes",
                 ic_params=None, discrete_real=None, g_disc=None, seed=0, stack_capacity=64,
                 simulation_type="cmm_euler_2d", forward_map=False, precision="fp64"):
        import numpy as np
        tin = TIME_INTEGRATION[time_integration]
        if l_order is None:
            l_order = {10: 1, 20: 2, 21: 2, 30: 3, 31: 3, 40: 4, 41: 4}[tin]
        p = EulerParams()
        p.g_coarse, p.g_fine, p.g_psi, p.g_vort = g_coarse, g_fine, g_psi, g_vort
        p.dt = dt
        p.time_integration_num = tin
        p.map_update_order_num = MAP_UPDATE_ORDER[map_update_order]
        p.l_order = l_order
        p.lagrange_init_higher_order = int(lagrange_init_higher_order)
        p.map_epsilon, p.incomp_threshold, p.freq_cut_psi = map_epsilon, incomp_threshold, freq_cut_psi
        p.molly_stencil, p.skip_remapping = molly_stencil, int(skip_remapping)
        p.ic = IC[initial_condition]
        for i, v in enumerate(ic_params or []):
            p.ic_params[i] = v
        self._keep = []
        if initial_condition == "random_spectrum" or discrete_real is not None:
            p.discrete = 1
            p.g_disc = g_disc if g_disc is not None else g_fine
            if discrete_real is not None:
                d = np.ascontiguousarray(discrete_real, dtype=np.float64)
                self._keep.append(d)
                p.disc_real = d.ctypes.data_as(DP)
        p.seed = seed
        p.stack_capacity = stack_capacity
        p.sim = SIM[simulation_type]
        p.forward_map = int(forward_map)
        p.precision = PRECISION[precision]
        self.params = p
        self.l_order = l_order
        h = C.c_void_p()
        _check(lib().cmm_euler_create(C.byref(p), C.byref(h)), "cmm_euler_create")
        self.h = h
        self.sets = []

    def close(self):
        if self.h:
            lib().cmm_euler_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_particles(self, pos, scheme=30, init_time=0.0, tau=0.0, init_vel=False, vel=None):
        """A particle set: fluid particles (tau 0) or inertial ones (tau != 0), moving from
        init_time on (cmm_euler_add_particle_set)."""
        import numpy as np
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        s = C.c_int()
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        _check(lib().cmm_euler_add_particle_set(self.h, pos.size // 2, scheme, float(init_time), float(tau),
                                                int(init_vel), pos.ctypes.data_as(DP),
                                                None if v is None else v.ctypes.data_as(DP), C.byref(s)),
               "cmm_euler_add_particle_set")
        self.sets.append(pos.size // 2)
        return s.value

    def particle_velocities(self, s):
        """inertial sets: their velocities; fluid sets: (d psi/dx, d psi/dy) at the positions"""
        import numpy as np
        out = np.zeros(2 * self.sets[s])
        _check(lib().cmm_euler_get_particle_velocities(self.h, s, out.ctypes.data_as(DP)),
               "cmm_euler_get_particle_velocities")
        return out

    def init(self):
        w = (C.c_double * 8)()
        n = C.c_int()
        _check(lib().cmm_euler_init(self.h, C.cast(w, DP), C.byref(n)), "cmm_euler_init")
        return [w[i] for i in range(n.value)]

    def step(self, dt_now):
        info = StepInfo()
        _check(lib().cmm_euler_step(self.h, dt_now, C.byref(info)), "cmm_euler_step")
        return info

    def field(self, name):
        import numpy as np
        p = self.params
        size = {"chix": 4 * p.g_coarse.n, "chiy": 4 * p.g_coarse.n, "chix_f": 4 * p.g_coarse.n,
                "chiy_f": 4 * p.g_coarse.n, "psi": 4 * p.g_psi.n * self.l_order,
                "w_fine": 4 * p.g_fine.n, "w_coarse": p.g_coarse.n, "disc": 4 * p.g_disc.n}[name]
        out = np.zeros(size)
        _check(lib().cmm_euler_get_field(self.h, FIELD[name], out.ctypes.data_as(DP)), "cmm_euler_get_field")
        return out

    def particles(self, s):
        import numpy as np
        out = np.zeros(2 * self.sets[s])
        _check(lib().cmm_euler_get_particles(self.h, s, out.ctypes.data_as(DP)), "cmm_euler_get_particles")
        return out

    def diagnostics(self):
        """energy, enstrophy, palinstrophy, max |omega| (as the reference's Mesure files)"""
        out = (C.c_double * 4)()
        _check(lib().cmm_euler_diagnostics(self.h, C.cast(out, DP)), "cmm_euler_diagnostics")
        return list(out)

    def sample(self, g_sample, scalar_ic=-1):
        """Fields through the whole map stack on g_sample (cmm_euler_sample): a dict of arrays
        chix, chiy, w, theta, lap_w, gradx_w, grady_w, psi, ux, uy."""
        import numpy as np
        n = g_sample.nx * g_sample.ny
        out = np.zeros(10 * n)
        _check(lib().cmm_euler_sample(self.h, g_sample, scalar_ic, out.ctypes.data_as(DP)), "cmm_euler_sample")
        names = ("chix", "chiy", "w", "theta", "lap_w", "gradx_w", "grady_w", "psi", "ux", "uy")
        return {k: out[i * n:(i + 1) * n] for i, k in enumerate(names)}

    def vp_sample(self, g_sample):
        """Vlasov-Poisson sample outputs on g_sample (cmm_euler_vp_sample): positions x, v through
        the whole stack, f there, the sample's electric field e (dphi/dx, nx values) and the
        measures etot, ekin, epot, mass, momentum, emodes (4)."""
        import numpy as np
        n, nx = g_sample.nx * g_sample.ny, g_sample.nx
        out = np.zeros(3 * n + nx + 9)
        _check(lib().cmm_euler_vp_sample(self.h, g_sample, out.ctypes.data_as(DP)), "cmm_euler_vp_sample")
        m = out[3 * n + nx:]
        return {"x": out[:n], "v": out[n:2 * n], "f": out[2 * n:3 * n], "e": out[3 * n:3 * n + nx], "etot": m[0],
                "ekin": m[1], "epot": m[2], "mass": m[3], "momentum": m[4], "emodes": m[5:9].copy()}

    def zoom(self, g_zoom, ics=(), want_psi=True):
        """A zoom window (cmm_euler_zoom): positions x, y through the whole stack on g_zoom, the
        initial conditions ics there (-1: the run's own), and psi at the window's points."""
        import numpy as np
        n = g_zoom.nx * g_zoom.ny
        k = len(ics)
        out = np.zeros((3 + k) * n)
        arr = (C.c_int * max(k, 1))(*ics)
        _check(lib().cmm_euler_zoom(self.h, g_zoom, arr, k, 1 if want_psi else 0, out.ctypes.data_as(DP)),
               "cmm_euler_zoom")
        res = {"x": out[:n], "y": out[n:2 * n], "ic": [out[(2 + i) * n:(3 + i) * n] for i in range(k)]}
        if want_psi:
            res["psi"] = out[(2 + k) * n:(3 + k) * n]
        return res

    def restart(self, t_restart, stack_maps, chix, chiy, psi_old):
        """Restart from a saved state (cmm_euler_restart): stack_maps is a list of (chix, chiy) host
        arrays (oldest first), chix/chiy the active map, psi_old the older psi levels."""
        import numpy as np
        st = np.ascontiguousarray(np.concatenate([np.concatenate([x, y]) for x, y in stack_maps])
                                  if stack_maps else np.zeros(1))
        cx, cy = np.ascontiguousarray(chix, dtype=np.float64), np.ascontiguousarray(chiy, dtype=np.float64)
        po = np.ascontiguousarray(np.concatenate(list(psi_old)) if len(psi_old) else np.zeros(1))
        _check(lib().cmm_euler_restart(self.h, float(t_restart), len(stack_maps), st.ctypes.data_as(DP),
                                       cx.ctypes.data_as(DP), cy.ctypes.data_as(DP), po.ctypes.data_as(DP)),
               "cmm_euler_restart")

    def stack_map(self, m):
        """Map m of the stack (0 = oldest) as host arrays (chix, chiy)."""
        import numpy as np
        n = 4 * self.params.g_coarse.n
        x, y = np.zeros(n), np.zeros(n)
        _check(lib().cmm_euler_get_stack_map(self.h, m, x.ctypes.data_as(DP), y.ctypes.data_as(DP)),
               "cmm_euler_get_stack_map")
        return x, y

    def forward_stack_map(self, m):
        """Forward map m of the forward stack (0 = oldest) as host arrays (chix, chiy)."""
        import numpy as np
        n = 4 * self.params.g_coarse.n
        x, y = np.zeros(n), np.zeros(n)
        _check(lib().cmm_euler_get_forward_stack_map(self.h, m, x.ctypes.data_as(DP), y.ctypes.data_as(DP)),
               "cmm_euler_get_forward_stack_map")
        return x, y

    def restart_forward(self, stack_maps, chix, chiy):
        """The forward part of a restart (cmm_euler_restart_forward), after restart()."""
        import numpy as np
        st = np.ascontiguousarray(np.concatenate([np.concatenate([x, y]) for x, y in stack_maps])
                                  if stack_maps else np.zeros(1))
        cx, cy = np.ascontiguousarray(chix, dtype=np.float64), np.ascontiguousarray(chiy, dtype=np.float64)
        _check(lib().cmm_euler_restart_forward(self.h, len(stack_maps), st.ctypes.data_as(DP), cx.ctypes.data_as(DP),
                                               cy.ctypes.data_as(DP)), "cmm_euler_restart_forward")

    def request_remap(self):
        """The next step remaps whatever E_inc is (forwarded-particle start)."""
        _check(lib().cmm_euler_request_remap(self.h), "cmm_euler_request_remap")

    def forward_positions(self, g_target):
        """The forward map through the forward stack on g_target: (x, y) host arrays."""
        import numpy as np
        n = g_target.nx * g_target.ny
        out = np.zeros(2 * n)
        _check(lib().cmm_euler_forward_positions(self.h, g_target, out.ctypes.data_as(DP)),
               "cmm_euler_forward_positions")
        return out[:n], out[n:]

    def forward_points(self, g_bounds, pos, first_map=0, apply_current=True):
        """Forwarded particles (x, y interleaved) through forward maps first_map.. and the current one."""
        import numpy as np
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros_like(pos)
        _check(lib().cmm_euler_forward_points(self.h, g_bounds, pos.size // 2, pos.ctypes.data_as(DP), first_map,
                                              int(apply_current), out.ctypes.data_as(DP)), "cmm_euler_forward_points")
        return out

    def vp_diagnostics(self):
        """Vlasov-Poisson: Etot, Ekin, Epot, Mass, Momentum, Emode1..4"""
        out = (C.c_double * 9)()
        _check(lib().cmm_euler_vp_diagnostics(self.h, C.cast(out, DP)), "cmm_euler_vp_diagnostics")
        return list(out)

    def state(self):
        st, t, ns = C.c_int(), C.c_double(), C.c_int()
        _check(lib().cmm_euler_state(self.h, C.byref(st), C.byref(t), C.byref(ns)), "cmm_euler_state")
        return st.value, t.value, ns.value

    def set_check_tile(self, tile, max_blocks=8192):
        """the check's tile edge and block cap, as the autotuner chooses them (E_inc does not
        depend on the cap; tile edges round differently near 1e-14, and grouped batch steps need
        one tile edge in all engines)"""
        _check(lib().cmm_euler_set_check_tile(self.h, int(tile), int(max_blocks)), "cmm_euler_set_check_tile")

    def set_check_fusion(self, on):
        """the step's incompressibility check in the psi sampler's pass (on by default where possible)"""
        _check(lib().cmm_euler_set_check_fusion(self.h, int(on)), "cmm_euler_set_check_fusion")

    def set_psi_precombine(self, on):
        """the psi solve also writes the next step's combined psi fields (on by default)"""
        _check(lib().cmm_euler_set_psi_precombine(self.h, int(on)), "cmm_euler_set_psi_precombine")

    def psi_precombine(self):
        on = C.c_int()
        _check(lib().cmm_euler_get_psi_precombine(self.h, C.byref(on)), "cmm_euler_get_psi_precombine")
        return bool(on.value)

    def check_fusion(self):
        on = C.c_int()
        _check(lib().cmm_euler_get_check_fusion(self.h, C.byref(on)), "cmm_euler_get_check_fusion")
        return bool(on.value)

    def profile(self, enable=True, every=1):
        """per-kernel events on every `every`-th step (enable False: none); clears the counters"""
        _check(lib().cmm_euler_profile(self.h, int(every) if enable else 0), "cmm_euler_profile")

    def profile_read(self):
        ms = (C.c_double * PROF_SLOTS)()
        cnt = (C.c_int * PROF_SLOTS)()
        _check(lib().cmm_euler_profile_read(self.h, C.cast(ms, DP), C.cast(cnt, C.POINTER(C.c_int))),
               "cmm_euler_profile_read")
        names = PROF_NAMES
        return {n: (ms[i], cnt[i]) for i, n in enumerate(names)}


def device_info():
    buf = C.create_string_buffer(512)
    _check(lib().cmm_device_info(buf, 512), "cmm_device_info")
    return buf.value.decode()


def _mix64(z):
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on numpy uint64 arrays."""
    import numpy as np
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
        return z ^ (z >> np.uint64(31))


def uniform_particles(n, seed, center=(TWO_PI / 2, TWO_PI / 2), width=(TWO_PI, TWO_PI), bounds=(0.0, TWO_PI, 0.0, TWO_PI)):
    """Positions of the 'uniform' particle initial condition: u in [0,1) from a counter-based
    SplitMix64 stream (u_j = mix64(mix64(seed) + j) >> 11 / 2^53, j = 2i for x and 2i+1 for y) in
    place of the reference's cuRAND generator, then the reference's rescale and periodic wrap
    (Cuda-CMM/src/numerical/cmm-particles.cu:30-46)."""
    import numpy as np
    j = np.arange(2 * n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix64((_mix64(np.array([seed], dtype=np.uint64))[0] + j).astype(np.uint64))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    pos = np.empty(2 * n)
    lx, ly = bounds[1] - bounds[0], bounds[3] - bounds[2]
    px = center[0] + width[0] * (u[0::2] - 0.5)
    py = center[1] + width[1] * (u[1::2] - 0.5)
    pos[0::2] = px - np.floor((px - bounds[0]) / lx) * lx
    pos[1::2] = py - np.floor((py - bounds[2]) / ly) * ly
    return pos


class Batch:
    """Engines stepped together (cmm_batch_*): one wait per step; with grouping (FP64 Vlasov-Poisson
    engines in their steady state) every kind of launch of all engines as one persistent launch,
    else with graphs the queued work of all engines replayed from one captured graph per step
    phase (a branch per engine)."""

    def __init__(self, engines, use_graphs=True, grouping=False):
        self.engines = list(engines)
        arr = (VP * len(self.engines))(*[e.h.value for e in self.engines])
        h = C.c_void_p()
        _check(lib().cmm_batch_create(arr, len(self.engines), int(use_graphs), C.byref(h)), "cmm_batch_create")
        self.h = h
        _check(lib().cmm_batch_set_grouping(self.h, int(grouping)), "cmm_batch_set_grouping")

    def step(self, dt):
        infos = (StepInfo * len(self.engines))()
        _check(lib().cmm_batch_step(self.h, dt, infos), "cmm_batch_step")
        return list(infos)

    MODES = {0: "eager", 1: "graphs", 2: "grouped", 3: "grouped_persistent", 4: "grouped_lanes"}

    def set_mode(self, name):
        """the launch form by name (MODES)"""
        code = {v: k for k, v in self.MODES.items()}[name]
        _check(lib().cmm_batch_set_mode(self.h, code), "cmm_batch_set_mode")

    def autotune(self, dt, steps_per_mode=4):
        """cmm_batch_autotune: the fastest launch form over real steps (the run advances)"""
        m, ms = C.c_int(), C.c_double()
        _check(lib().cmm_batch_autotune(self.h, dt, int(steps_per_mode), C.byref(m), C.byref(ms)), "cmm_batch_autotune")
        return {"mode": self.MODES.get(m.value, str(m.value)), "ms_per_step": ms.value}

    def stats(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().cmm_batch_stats(self.h, C.byref(a), C.byref(b), C.byref(c)), "cmm_batch_stats")
        g = C.c_int()
        _check(lib().cmm_batch_grouped_steps(self.h, C.byref(g)), "cmm_batch_grouped_steps")
        return {"graphs_captured": a.value, "graph_launches": b.value, "eager_steps": c.value, "grouped_steps": g.value}

    def close(self):
        if self.h:
            lib().cmm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(engine):
    """cmm_euler_autotune: the check's tile edge and block cap, timed on the engine's map"""
    t, b, ms = C.c_int(), C.c_int(), C.c_double()
    _check(lib().cmm_euler_autotune(engine.h, C.byref(t), C.byref(b), C.byref(ms)), "cmm_euler_autotune")
    return {"tile": t.value, "max_blocks": b.value, "ms": ms.value}


def ensemble_grid(n, nv=None):
    """(x, v) grid of the ensemble workload: x in [0, 4 pi), v in [-8, 8)"""
    return grid(n, nv, (0.0, 2 * TWO_PI, -8.0, 8.0))
