"""ctypes binding of libedge_amd.so (the C ABI declared in include/edge_amd.h).

The library is loaded from the package directory only; a missing library is
an error (there is no CPU fallback for the element update).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libedge_amd.so")

EDGE_OK = 0
EDGE_ERR_ARG = 1
EDGE_NO_ID16 = 0xFFFF
FA_INTERIOR = 0
FA_FREE_SURFACE = 101
FA_OUTFLOW = 105
FA_PERIODIC = 106
FA_RUPTURE = 201


class EdgeError(RuntimeError):
    pass


class BoxParams(C.Structure):
    _fields_ = [("n", C.c_int64 * 3), ("origin", C.c_double * 3), ("extent", C.c_double * 3),
                ("bc", C.c_int32 * 3), ("jitter", C.c_double), ("seed", C.c_uint64), ("block", C.c_int64)]


class MeshOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ["ve_crds", "el_ve", "el_fa", "fa_el", "el_fa_el", "f_id", "v_id", "fa_type", "normals",
                 "tangents", "areas", "in_dia", "volume"]]


class MeshIn(C.Structure):
    _fields_ = [("n_el", C.c_int64), ("n_fa", C.c_int64)] + [
        (n, C.c_void_p) for n in ["ve_crds", "el_ve", "el_fa", "fa_el", "el_fa_el", "f_id", "v_id", "fa_type",
                                  "normals", "tangents", "areas"]]


class ElConfig(C.Structure):
    _fields_ = [("order", C.c_int32), ("n_runs", C.c_int32), ("precision", C.c_int32), ("kernel", C.c_int32)]


class ViscoConfig(C.Structure):
    _fields_ = [("n_rms", C.c_int32), ("freq_cen", C.c_double), ("freq_rat", C.c_double)]


class RuptureConfig(C.Structure):
    _fields_ = [("n_sub_faces", C.c_int32), ("n_sub_cells", C.c_int32), ("n_runs", C.c_int32)]


class TetsIn(C.Structure):
    _fields_ = [("n_ve", C.c_int64), ("n_el", C.c_int64), ("ve_crds", C.c_void_p), ("el_ve", C.c_void_p),
                ("n_bf", C.c_int64), ("bf_ve", C.c_void_p), ("bf_type", C.c_void_p), ("default_bc", C.c_int32)]


# symbol -> (restype, argtypes)
_SIGS = {
    "edge_last_error": (C.c_char_p, []),
    "edge_version": (C.c_char_p, []),
    "edge_dg_sizes": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "edge_dg_matrices": (C.c_int, [C.c_int] + [C.c_void_p] * 6),
    "edge_dg_eval_basis": (C.c_int, [C.c_int, C.c_int64, C.c_void_p, C.c_void_p]),
    "edge_dg_quad_tet": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p]),
    "edge_mesh_box_sizes": (C.c_int, [C.POINTER(BoxParams), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64)]),
    "edge_mesh_box": (C.c_int, [C.POINTER(BoxParams), C.POINTER(MeshOut)]),
    "edge_el_create": (C.c_int, [C.POINTER(ElConfig), C.POINTER(MeshIn), C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_void_p)]),
    "edge_el_create_raw": (C.c_int, [C.POINTER(ElConfig), C.POINTER(MeshIn), C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_void_p)]),
    "edge_el_destroy": (C.c_int, [C.c_void_p]),
    "edge_el_get_setup": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_local": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    "edge_el_neigh": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_step": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                               C.c_void_p]),
    "edge_el_error_norms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_energy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_autotune": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "edge_el_step_a": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_void_p]),
    "edge_el_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int64)]),
    "edge_el_create_visco": (C.c_int, [C.POINTER(ElConfig), C.POINTER(ViscoConfig), C.POINTER(MeshIn)] +
                             [C.c_void_p] * 5 + [C.POINTER(C.c_void_p)]),
    "edge_el_local_a": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "edge_el_neigh_a": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "edge_el_get_setup_a": (C.c_int, [C.c_void_p] + [C.c_void_p] * 6),
    "edge_mesh_tets_sizes": (C.c_int, [C.POINTER(TetsIn), C.POINTER(C.c_int64)]),
    "edge_mesh_tets": (C.c_int, [C.POINTER(TetsIn), C.c_void_p]),
    "edge_el_set_lts": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_local_lts": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_int32] + [C.c_void_p] * 5),
    "edge_el_neigh_lts": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32] + [C.c_void_p] * 5),
    "edge_el_local_lts_a": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_int32] + [C.c_void_p] * 6),
    "edge_el_neigh_lts_a": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32] + [C.c_void_p] * 6),
    "edge_el_point_basis": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "edge_recv_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "edge_recv_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    "edge_recv_eval_a": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "edge_src_create": (C.c_int, [C.c_void_p, C.c_int64] + [C.c_void_p] * 8 + [C.POINTER(C.c_void_p)]),
    "edge_src_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "edge_points_destroy": (C.c_int, [C.c_void_p]),
    "edge_rupture_create": (C.c_int, [C.POINTER(RuptureConfig), C.c_int64] + [C.c_void_p] * 10 +
                            [C.POINTER(C.c_void_p)]),
    "edge_rupture_destroy": (C.c_int, [C.c_void_p]),
    "edge_rupture_net_updates": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double] + [C.c_void_p] * 5),
    "edge_rupture_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_dg_fault_tables": (C.c_int, [C.c_int, C.POINTER(C.c_int32), C.c_void_p, C.c_void_p]),
    "edge_fault_create": (C.c_int, [C.c_void_p, C.POINTER(MeshIn)] + [C.c_void_p] * 6 + [C.POINTER(C.c_void_p)]),
    "edge_fault_destroy": (C.c_int, [C.c_void_p]),
    "edge_fault_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_fault_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "edge_fault_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_set_fault": (C.c_int, [C.c_void_p, C.c_void_p]),
    "edge_el_set_variant": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "edge_el_kernel_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "edge_el_l2_norms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "edge_el_l2_norms_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    "edge_el_neigh_local": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "edge_el_set_taps": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
}

_lib = None


def lib():
    """Loads libedge_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EdgeError(f"{LIB_PATH} is missing: build it with `python -m edge_amd.build`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(rc):
    if rc != EDGE_OK:
        raise EdgeError(f"edge_amd error {rc}: {lib().edge_last_error().decode()}")


def ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def declared_symbols(header=None):
    """Function names declared in include/edge_amd.h."""
    import re
    header = header or os.path.join(os.path.dirname(HERE), "include", "edge_amd.h")
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(edge_[a-z0-9_]+)\s*\(", txt)))
