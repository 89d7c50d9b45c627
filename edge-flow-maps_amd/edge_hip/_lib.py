"""ctypes binding of libedge_hip.so (the C-ABI declared in include/edge_hip.h).

The library is loaded from this package directory only; there is no fallback:
if it is missing or fails to load, every caller gets a RuntimeError.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# EDGE_HIP_LIBRARY: path of an alternative build of libedge_hip.so (e.g. one
# compiled with other tuning knobs); default: the in-tree build.
LIB_PATH = os.environ.get("EDGE_HIP_LIBRARY") or os.path.join(_HERE, "libedge_hip.so")

_P, _I, _L, _F, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float, ctypes.c_double
_CODES = {"p": _P, "i": _I, "l": _L, "f": _F, "d": _D}

# name -> argument codes (every function returns int status unless listed in _RESTYPE)
SIGNATURES = {
    "edge_last_error": "",
    "edge_abi_version": "",
    "edge_stream_sync": "p",
    # tables / FD
    "edge_pack_hermite_table": "ipppppppppppiiifp",
    "edge_calc_fd3": "ipppppppppiiifp",
    "edge_pack_table_hi": "ppppiiifp",
    # marches
    "edge_march_edge": "ppppppppppppiiifffp",
    "edge_march_edge_table": "pppppppppppppiiifffp",
    "edge_march_ed8": "pppppppppppppppiiifffp",
    "edge_march_ge": "pppppppppiiiffp",
    "edge_march_ed4": "pppppppppiiifffp",
    "edge_advance_map_faces": "pppppppppiiiffp",
    "edge_advance_points": "ipppppppliiiffp",
    "edge_dye_advance": "pppplliiiffp",
    "edge_compose_points": "pppppliiifp",
    "edge_calc_ed4_garm": "ppppppppppppppplfp",
    "edge_calc_ed": "ippppppppppplfp",
    "edge_offset_points": "pplfffp",
    "edge_eval_points": "ipiiipppppiiiplfppp",
    "edge_reset_identity": "pppppppppppiiip",
    "edge_reset_faces": "ppppppiiifp",
    "edge_coords": "ipiiifp",
    # driver kernels
    "edge_map_u": "ppppppppppppiiifp",
    "edge_map_u_psi": "pppppppiiifp",
    "edge_split_psi_hermite": "pppppppiiifp",
    "edge_map_smoke": "ippipppiiifp",
    "edge_map_smoke_psi": "ippipiiifp",
    "edge_clamp_u": "pppppppppiiip",
    "edge_sub_clamp_u": "ppppppiiip",
    "edge_map_u_add_scale": "pppppppppppppppffiiifp",
    "edge_map_u_sub_clamp": "pppppppppppppppiiifp",
    "edge_map_smoke_add_scale": "ipppffipppiiifp",
    "edge_advect_u_sl": "ppppppiiiffp",
    "edge_max_speed": "ppppiiip",
    "edge_axpy": "pppflp",
    "edge_scale": "pfplp",
    "edge_axpy_scale": "pppfflp",
    "edge_copy": "pplp",
    "edge_fill": "pflp",
    "edge_vortex_ring": "fffpppppppiiiip",
    "edge_split_central_vector": "ppppiiip",
    "edge_get_central_vector": "ppppiiip",
    "edge_curl": "ppiiifp",
    "edge_vtk_frame": "ppppippiiifp",
    "edge_backtrack_faces": "ppppippppppiiifp",
    # projection
    "edge_mg_create": "iiiiipp",
    "edge_mg_destroy": "p",
    "edge_mg_set_solid": "ppp",
    "edge_poisson": "ppppppppiiddpipp",
    "edge_mg_solve": "pppiiddpipp",
    "edge_mg_vcycle": "pppp",
    "edge_mg_apply_A": "pppp",
    "edge_mg_dense_bottom": "p",
    "edge_mg_set_graph": "pi",
    "edge_mg_set_tiles": "pi",
    # learned warm start
    "edge_mg_rhs": "pppppp",
    "edge_conv3d_bf16_fwd": "pipipppiiiiip",
    "edge_conv3d_bf16_bwd_data": "pipiippiiip",
    "edge_conv3d_bf16_bwd_weight": "pipipipiiip",
    "edge_unet_pool": "ppiiiip",
    "edge_unet_pool_bwd": "ppiiiip",
    "edge_unet_relu_bwd": "ppppiiip",
    "edge_unet_input": "ppipppp",
    "edge_unet_head": "ppppiipp",
    "edge_unet_loss_head_bwd": "pppppiipppppp",
    "edge_unet_head_bwd": "ppppiipppp",
    "edge_unet_pack_weights": "piippp",
    "edge_adam": "ppppiffffffp",
    "edge_zero_f32": "plp",
    # particles
    "edge_lp_acc": "pppiiipppllffffp",
    "edge_lp_vel": "pppllfffffp",
    "edge_lp_vel_diff": "pppiiipppllffffp",
    "edge_lp_p2g": "ppppppllfffffiiip",
    "edge_lp_p2g_noalloc": "ppppppllfffffiiip",
    "edge_lp_pos": "ppllffffp",
    "edge_lp_viscous": "ppppppiiiffp",
    "edge_lp_max_speed": "plpp",
}

_RESTYPE = {"edge_last_error": ctypes.c_char_p}

# statuses that are results rather than failures
NOCONV = -4

_lib = None


def lib():
    """Load libedge_hip.so once; raise loudly when it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libedge_hip.so not found at {LIB_PATH}; build it with "
                "`python edge-flow-maps_amd/csrc/build.py` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, codes in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = [_CODES[c] for c in codes]
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error():
    msg = lib().edge_last_error()
    return msg.decode() if msg else ""


def call(name, *args, ok=(0,)):
    """Call an exported function and raise RuntimeError on a status not in `ok`."""
    st = getattr(lib(), name)(*args)
    if st not in ok:
        raise RuntimeError(f"{name} failed (status {st}): {last_error()}")
    return st


def exported_symbols():
    return list(SIGNATURES)
