"""ctypes binding of libmeeuuw_hip.so (the C ABI declared in include/meeuuw_hip.h).

The library is mandatory: importing this module raises if the built .so is
missing, and every call raises if the GPU path fails.  There is no CPU
fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmeeuuw_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
SZ = ctypes.c_size_t
L = ctypes.c_long

# name -> (restype, argtypes)
SIGNATURES = {
    "mw_version": (I, []),
    "mw_last_error": (ctypes.c_char_p, []),
    "mw_local_to_global_V": (I, [I, P, P, P]),
    "mw_coo_stokes": (I, [I, P, P, I, P, P, P]),
    "mw_coo_K": (I, [I, P, P, P, P]),
    "mw_coo_G": (I, [I, P, P, P, P, P, P, P]),
    "mw_coo_scalar": (I, [I, I, P, P, P, P]),
    "mw_incidence_workspace_bytes": (SZ, [I]),
    "mw_incidence_build": (I, [I, I, P, I, P, P, P, P]),
    "mw_build_matrix_stokes": (I, [I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P,
                                   P, D, D, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mw_build_matrix_stokes_general": (I, [I, I, I, I] + [P] * 18
                                       + [D, D, P, P, P, P, I, I, I, P, I, P] + [P] * 15),
    "mw_build_matrix_energy": (I, [I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P,
                                   P, P, P, P, P, D, D, P, P, P, P, P, P]),
    "mw_build_velocity_mesh": (I, [I, I, I, D, D, D, D, D, P, P, P, P, P, P]),
    "mw_build_velocity_connectivity": (I, [I, I, I, P, P, P, P, P]),
    "mw_build_pressure_mesh": (I, [I, I, I, D, D, D, D, D, P, P, P, P, P, P]),
    "mw_build_pressure_connectivity": (I, [I, I, I, P, P]),
    "mw_locate_particles_box": (I, [I, P, P, P, D, D, P, P, P, I, I, P, P, P, P, P]),
    "mw_interpolate_fields": (I, [I, I, P, P, P, P, P, P, I, P, P]),
    "mw_advect_particles_box": (I, [I, D, I, P, P, P, P, P, D, D, D, D, I, I, P, P, P, P, P, P,
                                    P]),
    "mw_locate_particles_annulus": (I, [I, P, P, P, D, D, P, P, P, I, I, D, P, P, P, P, P]),
    "mw_advect_particles_curved": (I, [I, I, D, I, P, P, P, P, P, P, P, D, D, D, D, I, I, P, P, P,
                                       P, P, P, P]),
    "mw_interpolate_vel_on_pts": (I, [I, I, P, P, P, P, D, D, P, P, D, D, P, P, D, I, I, P, P, P,
                                      P, P, P]),
    "mw_ls_coefficients": (I, [I, I, P, P, I, P, P, P, P, P, P, P, P, P, P]),
    "mw_ls_workspace_bytes": (SZ, [I, I]),
    "mw_limiter": (I, [I, P, P, P, P, P, D, P]),
    "mw_nodal_gradients": (I, [I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mw_elemental_strain_rate": (I, [I, P, P, P, P, P, P, P, P, P]),
    "mw_sort_workspace_bytes": (SZ, [I, I]),
    "mw_sort_particles_by_element": (I, [I, I, P, P, P, P, P, P]),
    "mw_sort_particles_by_element_noalloc": (I, [I, I, P, P, P, P, P]),
    "mw_project_particles_on_elements": (I, [I, P, P, P, P, I, P, P, P, P]),
    "mw_project_particle_fields_on_nodes": (I, [I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mw_nodal_to_qpoints": (I, [I, P, I, I, P, P, P, P]),
    "mw_pressure_offset": (I, [I, P, I, I, P, P, P, P, D, P, P, P]),
    "mw_pressure_offset_axisymmetric": (I, [I, P, I, P, P, P, P, P, P, P, P, P]),
    "mw_nodal_pressure": (I, [I, I, P, P, P, P, P, P, P, P]),
    "mw_nodal_heat_flux": (I, [I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mw_global_quantities": (I, [I, P, P, P, P, P, P, P, P, P, D, P, P, P, P, P, P]),
    "mw_compute_Nu": (I, [D, D, P, P, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "mw_sim_create": (I, [P, P, P]),
    "mw_sim_destroy": (I, [P]),
    "mw_sim_set_temperature": (I, [P, P]),
    "mw_sim_set_swarm_reduced": (I, [P, P, P]),
    "mw_sim_set_swarm_material": (I, [P, P]),
    "mw_sim_set_swarm": (I, [P, L, P, P, P, P, P, P]),
    "mw_sim_step": (I, [P, P]),
    "mw_sim_field_size": (L, [P, I]),
    "mw_sim_get_field": (I, [P, I, P, L]),
    "mw_sim_operator": (I, [P, I, P, P]),
    "mw_sim_stage_timing": (I, [P, I]),
    "mw_sim_stage_times": (I, [P, P]),
    "mw_output_create": (I, [P, P, I, I, P, P]),
    "mw_output_step": (I, [P, P, I]),
    "mw_output_flush": (I, [P]),
    "mw_output_destroy": (I, [P]),
    "mw_batch_create": (I, [I, P, I, P]),
    "mw_batch_sim": (I, [P, I, P]),
    "mw_batch_step": (I, [P, I, P, P]),
    "mw_batch_destroy": (I, [P]),
    "mw_sim_time_kernel": (I, [P, I, I, P]),
    "mw_cnn_create": (I, [I, I, I, P]),
    "mw_cnn_destroy": (I, [P]),
    "mw_cnn_set_params": (I, [P, P, P]),
    "mw_cnn_get_params": (I, [P, P, P]),
    "mw_cnn_get_grads": (I, [P, P, P]),
    "mw_cnn_forward": (I, [P, I, P, P, P]),
    "mw_cnn_backward": (I, [P, I, P, P]),
    "mw_cnn_adam": (I, [P, D, P]),
    "mw_profile_enable": (I, [I]),
    "mw_uzawa3_solver_L2": (I, [I, I] + [P] * 15 + [P, P, P, P, D, P, P, P, P, P]),
    "mw_cd_u_eta": (I, [I, I] + [P] * 18 + [P, P, P, P, D, P, P, P, P, P, P]),
    "mw_profile_read": (I, [P, I]),
}

MW_EINVAL, MW_EHIP, MW_EOUTSIDE, MW_EEMPTY, MW_ENOCONV, MW_ERANGE, MW_EUNSUPPORTED = 1, 2, 3, 4, 5, 6, 7


class MeeuuwHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[mw status {code}] {msg}")
        self.code = code


def _load():
    # PyTorch-ROCm bundles its own libamdhip64 (soname libamdhip64.so.7) and its
    # libraries look it up by the unversioned name.  Loading torch first makes
    # this library bind to that same runtime; the other order would put two HIP
    # runtimes in one process, and the second one finds no device.
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m meeuuw_amd.build` "
            "(the package has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    """Map a status code to the reference's exception types."""
    if rc == 0:
        return
    msg = lib.mw_last_error().decode(errors="replace")
    if rc == MW_EINVAL:
        raise ValueError(msg)
    if rc in (MW_EOUTSIDE, MW_EEMPTY, MW_ERANGE):
        raise SystemExit(msg)  # the reference aborts with exit("...") in these cases
    raise MeeuuwHipError(rc, msg)


def call(name, *args):
    check(getattr(lib, name)(*args))


def ptr_array(ptrs, ctype=ctypes.c_void_p):
    """Host array of device pointers (for the *_host pointer-list parameters)."""
    arr = (ctype * len(ptrs))(*ptrs)
    return ctypes.cast(arr, ctypes.c_void_p), arr
