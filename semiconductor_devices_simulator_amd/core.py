"""Loader and ctypes bindings of the api3_* C ABI.

The product is the HIP-backed shared library ``libsemisim_hip.so`` built from
``csrc/`` by :mod:`semiconductor_devices_simulator_amd.build`.  It exports the
same C entry points as the reference OpenMP core (semiconductor_3d.c:4567-5353),
so the reference's ctypes front end can load it unchanged through its
``SEMISIM_SO`` hook (semibuild.py:516-522, sentinel symbol ``api3_set_caxis``
semibuild.py:28).  This module mirrors that loader and the argument/return
types the reference binds (main.py:172-217, 3494-3513).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libsemisim_hip.so"
LIB_PATH = os.path.join(HERE, "lib", LIB_NAME)
SENTINEL = "api3_set_caxis"          # semibuild.py:28

D = ctypes.c_double
I = ctypes.c_int
LL = ctypes.c_longlong
DP = ctypes.POINTER(D)
IP = ctypes.POINTER(I)
FP = ctypes.POINTER(ctypes.c_float)

# (name, restype, argtypes): the reference's hard and soft bindings
# (main.py:172-217, 3485, 3494-3513; validate.py:31-49 for the test hooks).
SIGNATURES = [
    ("api3_init", None, []),
    ("api3_add_material", I, [I, D]),
    ("api3_set_gate_stack", I, [I, I, IP, DP]),
    ("api3_get_ntp", I, [I]),
    ("api3_tun_eval", D, [I, I, IP, DP, D, D, D, D, D, DP]),
    ("api3_mlda_factor", D, [I, D, D, I]),
    ("api3_lombardi", D, [I, D, D, D, D, I]),
    ("api3_set_grid", None, [I, I, I, D, D, D]),
    ("api3_set_grid_coords", I, [I, I, I, DP, DP, DP]),
    ("api3_set_solver", None, [I, I, I, D, D, D, D]),
    ("api3_set_region", None, [I, I, I, I, I, I, I, D, D]),
    ("api3_add_contact", None, [I, I, I, I, I, I, D, D]),
    ("api3_add_contact_ex", I, [I, I, I, I, I, I, I, I, D, D, D, D]),
    ("api3_clear_contacts", None, []),
    ("api3_add_sheet_charge", None, [I, I, I, I, I, I, D]),
    ("api3_add_fixed_charge", None, [I, I, I, I, I, I, D]),
    ("api3_charge_partition", I, [IP]),
    ("api3_terminal_charges", I, [I, DP]),
    ("api3_charge_owner", I, [IP]),
    ("api3_set_models", None, [I]),
    ("api3_set_caxis", None, [I]),
    ("api3_get_caxis", I, []),
    ("api3_solve", I, []),
    ("api3_resolve", I, []),
    ("api3_set_contact_V", None, [I, D]),
    ("api3_get_contact_I", D, [I]),
    ("api3_get_contact_A", D, [I]),
    ("api3_get_contact_In", D, [I]),
    ("api3_get_contact_Ifloor", D, [I]),
    ("api3_get_contact_Iraw", D, [I]),
    ("api3_get_contact_Vsolved", D, [I]),
    ("api3_get_stat", LL, [I]),
    ("api3_get_time", D, [I]),
    ("api3_get_Nx", I, []), ("api3_get_Ny", I, []), ("api3_get_Nz", I, []),
    ("api3_get_conv", I, []), ("api3_get_iter", I, []), ("api3_get_res", D, []),
    ("api3_get_xs", D, [I]), ("api3_get_ys", D, [I]), ("api3_get_zs", D, [I]),
    ("api3_mat_ni", D, [I]), ("api3_mat_Eg", D, [I]),
    ("api3_bulk_copy", None, [I, DP]),
    ("api3_get_Emax", D, []),
    ("api3_face_J", I, [I, I, DP]),
    ("api3_iv_sweep", None, [I, D, D, I, DP, DP, IP]),
    ("api3_iv_sweep2", None, [I, D, D, I, DP, DP, DP, IP]),
    ("api3_get_ncon", I, []),
    ("api3_request_stop", None, []),
    ("api3_clear_stop", None, []),
    ("api3_stop_requested", I, []),
    ("api3_get_progress", None, [DP]),
    ("api3_set_threads", None, [I]),
    ("api3_set_verbose", None, [I]),
    ("api3_get_aa_depth", I, []),
    ("api3_set_node_data", None, [IP, DP, DP, DP]),
    ("api3_set_con_nodes", I, [I, I, IP, DP]),
    ("api3_set_mesh_graph", I, [I, I, IP, IP, DP, DP, DP, DP]),
    ("api3_is_graph", I, []),
    ("api3_set_walls", None, [DP]),
    ("api3_set_mat_tau", I, [I, D, D]),
    ("api3_carrier_sums", None, [DP]),
    ("api3_ins_field_max", D, [DP]),
    ("api3_get_state", I, [DP, DP, DP]),
    ("api3_set_state", I, [DP, DP, DP]),
    ("api3_set_predictor", None, [I]),
    ("api3_get_Cref", D, []),
]

# Entry points of the HIP library with no counterpart in the reference
# (precision mode, device info, kernel-level test hooks): bound when present.
EXT_SIGNATURES = [
    ("sdx_set_precision", I, [I]),
    ("sdx_get_precision", I, []),
    ("sdx_device_name", I, [ctypes.c_char_p, I]),
    ("sdx_kernel_stats", I, [DP, I]),
    ("sdx_reset_stats", None, []),
    ("sdx_microbench", D, [I, I]),
    ("sdx_tune", None, [I, I]),
    ("sdx_ktimer_enable", None, [I, I]),
    ("sdx_ktimer_read", I, [DP]),
    ("sdx_ktimer_pause", None, [I]),
    ("sdx_ws_mode", None, [I]),
    ("sdx_ws_stats", None, [DP]),
    ("sdx_ws_reset_stats", None, []),
    ("sdx_float_stats", None, [DP]),
    ("sdx_float_reset_stats", None, []),
    ("sdx_ws_config", None, [I, D]),
    ("sdx_ws_accept", None, [D]),
    ("sdx_nn_init", I, [I, I, I, I, I, ctypes.c_uint]),
    ("sdx_nn_free", None, []),
    ("sdx_nn_nparams", I, []),
    ("sdx_nn_get_params", I, [FP, I]),
    ("sdx_nn_set_params", I, [FP, I]),
    ("sdx_nn_set_scale", None, [D]),
    ("sdx_nn_get_scale", D, []),
    ("sdx_nn_forward_host", I, [FP, FP]),
    ("sdx_nn_grad_host", D, [FP, FP, FP]),
    ("sdx_nn_dataset_alloc", I, [I]),
    ("sdx_nn_dataset_size", I, []),
    ("sdx_nn_dataset_get", I, [I, FP, FP]),
    ("sdx_nn_dataset_put", I, [FP, FP]),
    ("sdx_nn_train", D, [I, I, D, ctypes.c_uint, FP]),
    ("sdx_nn_eval", D, [I, I]),
    ("sdx_test_prepare", I, []),
    ("sdx_test_set_state", None, [DP, DP, DP]),
    ("sdx_test_get_node", None, [I, DP]),
    ("sdx_test_poisson", D, [I, DP, DP, DP, DP]),
    ("sdx_test_upd_mob", None, [DP]),
    ("sdx_test_set_fmu", None, [DP]),
    ("sdx_test_cont", None, [I, DP, DP, DP, DP]),
    ("sdx_test_load_op", None, [DP, I, DP]),
    ("sdx_test_graph_op", I, [DP, DP]),
    ("sdx_test_graph_load_op", I, [DP, DP, I, DP]),
    ("sdx_test_graph_gmu", I, [DP, I]),
    ("sdx_test_graph_levels", I, []),
    ("sdx_test_matvec", None, [DP, DP]),
    ("sdx_test_mg_setup", I, []),
    ("sdx_test_level_dims", None, [I, IP]),
    ("sdx_test_level_op", None, [I, DP]),
    ("sdx_test_precond", None, [DP, DP]),
    ("sdx_test_fill_mg", None, [D]),
    ("sdx_test_pcg", I, [DP, DP, I, D]),
    ("sdx_test_bicgstab", D, [DP, DP, I, D, D, IP]),
    ("sdx_test_currents", None, [DP, DP, DP, DP]),
    ("sdx_test_cont_update", D, [I, DP, DP, DP, D, D]),
    ("sdx_test_ii_update", None, [DP]),
    ("sdx_test_aa_step", None, [DP, DP, I]),
    ("sdx_test_tun_paths", I, [IP, IP, IP, IP, IP, DP, IP, DP, DP]),
    ("sdx_test_tun", I, [DP, I, DP, DP, DP, DP]),
    ("sdx_test_csr", LL, [IP, IP, DP, LL]),
    ("sdx_dev_alloc", ctypes.c_void_p, [LL]),
    ("sdx_dev_free", None, [ctypes.c_void_p]),
    ("sdx_dev_copy", I, [ctypes.c_void_p, ctypes.c_void_p, LL, I]),
    ("sdx_csr_from_stencil", LL, [I, I, I, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, LL]),
    ("sdx_csr_from_stencil_noalloc", LL, [I, I, I, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, LL]),
    ("sdx_test_maj_flags", I, [I, DP, DP, ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(ctypes.c_ubyte)]),
]


class CoreError(RuntimeError):
    pass


def bind(lib, extra=True):
    """Attach restype/argtypes for every api3_* symbol the library exports."""
    for name, res, args in SIGNATURES + (EXT_SIGNATURES if extra else []):
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = list(args)
    return lib


def load_core(path=None):
    """The HIP core as a bound ctypes.CDLL (SEMISIM_SO overrides the path,
    semibuild.py:516-522).  Raises CoreError when the library is missing:
    there is no CPU fallback."""
    path = path or os.environ.get("SEMISIM_SO") or LIB_PATH
    if not os.path.isfile(path):
        raise CoreError(f"HIP solver core not built: {path} is missing "
                        "(run __graft_entry__.build() or python -m "
                        "semiconductor_devices_simulator_amd.build)")
    lib = ctypes.CDLL(path)
    if not hasattr(lib, SENTINEL):
        raise CoreError(f"{path} does not export {SENTINEL}: not an api3 solver core")
    return bind(lib)


def dptr(a):
    return a.ctypes.data_as(DP)


def iptr(a):
    return a.ctypes.data_as(IP)


def fptr(a):
    return a.ctypes.data_as(FP)
