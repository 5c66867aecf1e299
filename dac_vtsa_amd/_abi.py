"""ctypes binding of libvtsa (include/vtsa.h).

Loads the in-tree ``libvtsa.so`` built by ``dac_vtsa_amd/build.py``.  There is no fallback: if
the library is missing or the device is not usable, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvtsa.so")

D = C.c_double
I32 = C.c_int32
P = C.c_void_p
# vtsa_cons_shift mode bits (include/vtsa.h)
SHIFT_TRUNCATE, SHIFT_AT_ZERO, SHIFT_NO_REPORT = 1, 2, 4
OPT_GS_WHOLE_LINES, OPT_GS_FORWARD, OPT_GS_TWIST = 1, 2, 3   # vtsa_set_option


class ModelDesc(C.Structure):
    """Mirror of ``vtsa_model_desc``."""
    _fields_ = [
        ("Nr", I32), ("Nth", I32), ("Nz", I32),
        ("wall_on", I32), ("lumped", I32), ("hx_on", I32), ("jacket_on", I32),
        ("limiter", I32), ("disp_model", I32), ("coads_model", I32), ("water_basis_eq", I32),
        ("bed_model", I32), ("wall_model", I32),
        ("arrhenius", I32), ("mn_dynamic", I32), ("vac_krish", I32), ("pump_water", I32),
        ("end_kind", I32 * 2), ("end_check", I32 * 2), ("end_valve", I32 * 2), ("end_is_vacuum", I32 * 2),
        ("end_is_feed", I32 * 2), ("pad_", I32),
        ("end_P_target", D * 2), ("end_T_ext", D * 2), ("end_speed", D * 2), ("end_P_start", D * 2),
        ("end_y_ext", (D * 3) * 2),
        ("dz", D), ("dth", D), ("A_wall_in", D), ("A_wall_out", D), ("A_wall_axial", D), ("R_wall_mid", D),
        ("wall_thickness", D), ("half_dr_outer", D), ("C_wall", D),
        ("R", D), ("eps_b", D), ("eps_t", D), ("rho_b", D), ("cp_s", D), ("T_ref", D),
        ("cpg", D * 3), ("cpa", D * 2), ("dHa", D * 2), ("M", D * 3),
        ("mu_f1", D * 3), ("mu_a1", D * 3), ("mu_b1", D * 3), ("mu_f2", D * 3), ("mu_a2", D * 3), ("mu_b2", D * 3),
        ("k_c1", D * 3), ("k_n1", D * 3), ("k_c2", D * 3), ("k_n2", D * 3),
        ("mu_two", I32 * 3), ("pad2_", I32),
        ("wilke_mq", (D * 3) * 3), ("wilke_den", (D * 3) * 3),
        ("fuller_sqrtM_air", D), ("fuller_s_air", D), ("fuller_sqrtM_h2o", D), ("fuller_s_h2o", D),
        ("lam_const", D), ("lam_s", D), ("zs_B", D), ("zs_s", D),
        ("co2_q_inf0", D), ("co2_chi", D), ("co2_b0", D), ("co2_dH0", D), ("co2_tau0", D), ("co2_alpha", D),
        ("co2_T0", D), ("co2_p_reg", D),
        ("h2o_q_m", D), ("h2o_C", D), ("h2o_D", D), ("h2o_F", D), ("h2o_G", D), ("h2o_E10_slope", D),
        ("h2o_E10_intercept", D), ("h2o_rh_cap", D),
        ("wadst_b0", D), ("wadst_q_inf0", D), ("wadst_tau0", D), ("wadst_chi", D), ("wadst_alpha", D),
        ("wadst_dH", D), ("wadst_A", D),
        ("mech_f_blocked_max", D), ("mech_k", D), ("mech_phi_dry", D), ("mech_A", D), ("mech_dH_wet", D),
        ("mech_n", D),
        ("sb_gamma", D), ("sb_beta", D),
        ("k_ldf", D * 2), ("Ea", D * 2), ("k_T_ref", D),
        ("ka", D), ("kb", D), ("dp", D), ("er_g1", D), ("er_g2", D), ("er_beta", D), ("Pe_T", D), ("Pe_r", D),
        ("Pe_z", D), ("N_ratio", D), ("alpha_const", D), ("U_lumped", D),
        ("u_nom", D), ("tau_valve", D),
        ("h_jac", D), ("h_amb", D), ("T_amb", D), ("P_amb", D), ("T_jac_set", D), ("T_jac_prev", D),
        ("jac_tau", D), ("hA_v", D), ("lam_w", D),
        ("gam", D), ("eta_vac", D), ("eta_bl", D),
        ("typ", D * 7), ("atol", D * 7),
    ]


class GmresInfo(C.Structure):
    _fields_ = [("info", I32), ("iters", I32), ("breakdown", I32), ("pad_", I32),
                ("bnorm", D), ("presid", D), ("rnorm", D), ("atol_eff", D)]


_SIGS = {
    "vtsa_create": [C.POINTER(ModelDesc)] + [P] * 9 + [C.POINTER(P)],
    "vtsa_destroy": [P],
    "vtsa_reset": [P, C.POINTER(ModelDesc)] + [P] * 9 + [P],
    "vtsa_sizes": [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)],
    "vtsa_rhs": [P, D, P, P, P, P],
    "vtsa_stage_residual": [P, D, P, D, P, P, P],
    "vtsa_diag": [P, D, P, P, P, P, P, P, P, P],
    "vtsa_equilibrium": [P, I32, P, P, P, P, P, P, P],
    "vtsa_fd_jacobian": [P, D, P, P, P],
    "vtsa_jacobian_get": [P, P, P, P, P],
    "vtsa_precond_setup": [P, D, P, I32, I32, I32, P],
    "vtsa_precond_apply": [P, P, P, P],
    "vtsa_gmres": [P, D, P, P, P, D, D, P, P, P, D, D, I32, C.POINTER(GmresInfo), P],
    "vtsa_scale": [P, P, D, P, P],
    "vtsa_rms": [P, P, P, I32, C.POINTER(D), P],
    "vtsa_rms_noalloc": [P, P, I32, P, P],
    "vtsa_gmres_hessenberg": [P, P, C.POINTER(I32)],
    "vtsa_defl_set": [P, I32, I32, P, P, P],
    "vtsa_defl_apply": [P, P, P, P],
    "vtsa_newton_eval": [P, D, P, D, P, P, P, P, C.POINTER(D), C.POINTER(D), P],
    "vtsa_newton_residual": [P, D, P, D, P, P, P, P, P, P],
    "vtsa_newton_try": [P, D, P, P, D, D, P, P, P, P, P, C.POINTER(D), C.POINTER(D), C.POINTER(D), P],
    "vtsa_cons": [P, P, P, P, P, P],
    "vtsa_inventories": [P, P, P, P],
    "vtsa_state_roundtrip": [P, P, P],
    "vtsa_cons_norm2": [P, P, P, P],
    "vtsa_cons_shift": [P, P, P, P, I32, P, P],
    "vtsa_cons_stash": [P, P, P, I32, P],
    "vtsa_cons_fetch": [P, I32, P, P, P],
    "vtsa_vec_op": [P, I32, D, D, D, D, P, P, P, P, P, P],
    "vtsa_error_norm": [P, P, P, P, D, C.POINTER(D), P],
    "vtsa_device_count": [C.POINTER(I32)],
    "vtsa_device_alloc": [C.c_size_t, C.POINTER(P)],
    "vtsa_device_free": [P],
    "vtsa_copy_h2d": [P, P, C.c_size_t, P],
    "vtsa_copy_d2h": [P, P, C.c_size_t, P],
    "vtsa_copy_d2d": [P, P, C.c_size_t, P],
    "vtsa_fill": [P, D, C.c_size_t, P],
    "vtsa_sync": [P],
    "vtsa_stream_create": [C.POINTER(P)],
    "vtsa_stream_destroy": [P],
    "vtsa_time_rhs": [P, D, P, P, I32, C.POINTER(D), P],
    "vtsa_time_fdjac": [P, D, P, P, I32, C.POINTER(D), P],
    "vtsa_outlet_row": [P, D, P, I32, P, P],
    "vtsa_set_option": [P, I32, I32],
    "vtsa_profile": [P, I32],
    "vtsa_profile_read": [P, C.POINTER(C.c_int64), C.POINTER(D), C.POINTER(D)],
    "vtsa_mlp_sizes": [C.POINTER(I32)] * 5,
    "vtsa_mlp_create": [I32, C.POINTER(P)],
    "vtsa_mlp_destroy": [P],
    "vtsa_mlp_set_params": [P, P, P],
    "vtsa_mlp_get_params": [P, P, P],
    "vtsa_mlp_set_norm": [P, P, P],
    "vtsa_mlp_forward": [P, P, I32, P, P],
    "vtsa_mlp_train_step": [P, P, P, I32, C.c_float, C.POINTER(C.c_float), P, P],
    "vtsa_mlp_features": [P, P, P, P, P, D, D, P, P],
    "vtsa_mlp_targets": [P, P, P, P, P, P, P],
    "vtsa_mlp_apply": [P, P, P, P, P, P, P, P],
    "vtsa_version": [],
    "vtsa_last_error": [],
}
EXPORTS = tuple(_SIGS)

_lib = None


def lib():
    """The loaded library (raises if it is missing or cannot be loaded)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dac_vtsa_amd.build` "
                               "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name in ("vtsa_version", "vtsa_last_error") else C.c_int
        _lib = L
    return _lib


def call(name, *args):
    """Call an ABI function and raise on a negative (API/HIP) status; returns the status."""
    rc = getattr(lib(), name)(*args)
    if rc < 0:
        msg = lib().vtsa_last_error().decode()
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc
