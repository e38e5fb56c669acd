"""ctypes bindings of libcppimpact_hip.so (include/cppimpact_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
cppimpact_amd/csrc``).  Loading fails loudly when it is missing: there is no
CPU fallback for the step.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcppimpact_hip.so")

CPI_OK = 0
STATUS = {
    0: "CPI_OK", -1: "CPI_ERR_INVALID", -2: "CPI_ERR_HIP", -3: "CPI_ERR_ALLOC",
    -4: "CPI_ERR_MESH", -5: "CPI_ERR_NUMERIC", -6: "CPI_ERR_IO", -7: "CPI_ERR_STATE",
    -8: "CPI_ERR_NO_DEVICE",
}
PHYSICS_COMPAT, PHYSICS_EP = 0, 1
PREC_F64, PREC_F32 = 0, 1
SCATTER_BLOCK, SCATTER_COLOR, SCATTER_ATOMIC = 0, 1, 2
QUAD_DEFAULT, QUAD_REF5, QUAD_REF4, QUAD_EXACT4, QUAD_CENTROID = -1, 0, 1, 2, 3
MASS_DEFAULT, MASS_ROWSUM, MASS_HRZ = -1, 0, 1
RETURN_EXACT, RETURN_MLP = 0, 1
E_KINETIC, E_INTERNAL, E_PLASTIC, E_PENALTY, E_GRAVITY, E_CONTACTF, E_BALANCE, E_MINDT = range(8)
NENERGY = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class CpiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_elements", C.c_int32),
                ("nodes_per_element", C.c_int32), ("xloc", _dp), ("element_nodes", _ip),
                ("num_slave_nodes", C.c_int32), ("slave_nodes", _ip),
                ("num_fixed_nodes", C.c_int32), ("fixed_nodes", _ip)]


class Material(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("E", "rho", "nu", "beta", "H", "Y0")]


class WallDesc(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("dim", C.c_int32), ("normal", C.c_int32),
                ("location", C.c_double), ("stiffness", C.c_double)]


class Options(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "physics", "precision", "scatter", "quadrature", "mass", "stress_update",
        "use_graph", "track_energy", "contact_slave_only", "block_elements", "device",
        "fixed_bc", "debug_fill_nan", "stream_state")] + [("gravity", C.c_double * 3),
                                          ("mlp_tolerance", C.c_double)]


class Info(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "num_nodes", "num_elements", "nodes_per_element", "num_quadrature_pts", "num_colors",
        "num_blocks", "num_boundary_nodes", "num_partial_slots")] + [
        ("bytes_per_step", C.c_double), ("elem_kernel_bytes", C.c_double),
        ("node_kernel_bytes", C.c_double), ("dt", C.c_double), ("step", C.c_int64)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("total_ms", C.c_double), ("launches", C.c_int64)]


class BoxSpec(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("quadratic", C.c_int32),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("origin", C.c_double * 3), ("jitter", C.c_double), ("seed", C.c_uint64)]


class CylSpec(C.Structure):
    _fields_ = [("ntheta", C.c_int32), ("nz", C.c_int32), ("nr", C.c_int32),
                ("quadratic", C.c_int32), ("radius", C.c_double), ("length", C.c_double),
                ("thickness", C.c_double), ("origin", C.c_double * 3)]


class InpCounts(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "num_nodes", "num_elements", "num_node_sets", "num_set_indices", "num_fixed_nodes",
        "num_slave_nodes", "names_bytes")]


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float),
                ("lr_final", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("seed", C.c_uint64)]


MLP_INPUTS, MLP_OUTPUTS, MLP_NPARAMS, SAMPLE_FLOATS = 13, 7, 5511, 20
MLP_ELASTIC, MLP_ACCEPTED, MLP_FALLBACK = 0, 1, 2
_fp = C.POINTER(C.c_float)

# every exported symbol and its (restype, argtypes)
SIGNATURES = {
    "cpi_default_options": (None, [C.POINTER(Options)]),
    "cpi_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Material), C.POINTER(WallDesc),
                             C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "cpi_destroy": (None, [C.c_void_p]),
    "cpi_last_error": (C.c_char_p, [C.c_void_p]),
    "cpi_set_state": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cpi_stable_dt": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cpi_stable_dt_noalloc": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cpi_init": (C.c_int, [C.c_void_p, C.c_double]),
    "cpi_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "cpi_synchronize": (C.c_int, [C.c_void_p]),
    "cpi_get_state": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "cpi_get_increment": (C.c_int, [C.c_void_p, _dp]),
    "cpi_get_element_state": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cpi_get_stress_strain": (C.c_int, [C.c_void_p, _dp, _dp]),
    "cpi_get_energy": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int64)]),
    "cpi_get_energy_history": (C.c_int, [C.c_void_p, _dp, C.c_int32, C.POINTER(C.c_int32)]),
    "cpi_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "cpi_get_coloring": (C.c_int, [C.c_void_p, _ip, C.POINTER(C.c_int32)]),
    "cpi_time_steps": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "cpi_profile_kernels": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(KernelTime), C.c_int32,
                                      C.POINTER(C.c_int32)]),
    "cpi_debug_strain": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _dp, _dp]),
    "cpi_bandwidth_probe": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, _dp]),
    "cpi_greedy_coloring": (C.c_int, [C.c_int32, C.c_int32, _ip, C.c_int32, _ip]),
    "cpi_mesh_box": (C.c_int, [C.POINTER(BoxSpec), _ip, _ip, _dp, _ip]),
    "cpi_mesh_cylinder": (C.c_int, [C.POINTER(CylSpec), _ip, _ip, _dp, _ip]),
    "cpi_inp_read": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(InpCounts), _dp, _ip, _ip, _ip,
                               _ip, _ip, C.c_char_p]),
    "cpi_inp_write": (C.c_int, [C.c_char_p, C.POINTER(MeshDesc)]),
    "cpi_mlp_default_train_opts": (None, [C.POINTER(MlpTrainOpts)]),
    "cpi_mlp_create": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(Material),
                                 C.POINTER(C.c_void_p)]),
    "cpi_mlp_destroy": (None, [C.c_void_p]),
    "cpi_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "cpi_mlp_set_params": (C.c_int, [C.c_void_p, _fp]),
    "cpi_mlp_get_params": (C.c_int, [C.c_void_p, _fp]),
    "cpi_mlp_set_scaling": (C.c_int, [C.c_void_p, _fp, _fp, _fp, _fp]),
    "cpi_mlp_get_scaling": (C.c_int, [C.c_void_p, _fp, _fp, _fp, _fp]),
    "cpi_mlp_fit_scaling": (C.c_int, [C.c_void_p, _fp, C.c_int32]),
    "cpi_mlp_targets": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp, _fp, _ip]),
    "cpi_mlp_forward": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "cpi_mlp_gradient": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, _fp, _fp]),
    "cpi_mlp_train": (C.c_int, [C.c_void_p, _fp, C.c_int32, C.POINTER(MlpTrainOpts), _fp]),
    "cpi_mlp_return_map": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _fp, _fp, _fp, _fp,
                                     _fp, _ip]),
    "cpi_set_surrogate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cpi_get_surrogate_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]),
    "cpi_reset_surrogate_stats": (C.c_int, [C.c_void_p]),
    "cpi_samples_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "cpi_samples_subsample": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64]),
    "cpi_samples_get": (C.c_int, [C.c_void_p, _fp, C.c_int32, _ip]),
    "cpi_samples_reset": (C.c_int, [C.c_void_p]),
}

_LIB = None


def lib():
    """Loads the in-tree library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C cppimpact_amd/csrc); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def check(rc, ctx=None):
    if rc < 0:
        msg = lib().cpi_last_error(ctx)
        raise CpiError(rc, msg.decode() if msg else "")
    return rc


def check_mlp(rc, m=None):
    if rc < 0:
        msg = lib().cpi_mlp_last_error(m)
        raise CpiError(rc, msg.decode() if msg else "")
    return rc


def default_options(**kw):
    o = Options()
    lib().cpi_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "gravity":
            o.gravity = (C.c_double * 3)(*v)
        else:
            setattr(o, k, v)
    return o


def box_mesh(nx, ny, nz, lx, ly, lz, origin=(0.0, 0.0, 0.0), quadratic=False, jitter=0.0,
             seed=0):
    """Structured hex grid split into 6 tets per cell (see cpi_mesh_box)."""
    s = BoxSpec(nx, ny, nz, int(quadratic), lx, ly, lz, (C.c_double * 3)(*origin), jitter, seed)
    nn, ne = C.c_int32(), C.c_int32()
    check(lib().cpi_mesh_box(C.byref(s), C.byref(nn), C.byref(ne), None, None))
    npe = 10 if quadratic else 4
    x = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, npe), dtype=np.int32)
    check(lib().cpi_mesh_box(C.byref(s), C.byref(nn), C.byref(ne), dptr(x), iptr(conn)))
    return x, conn


def cylinder_mesh(ntheta, nz, nr, radius, length, thickness, origin=(0.0, 0.0, 0.0),
                  quadratic=False):
    s = CylSpec(ntheta, nz, nr, int(quadratic), radius, length, thickness,
                (C.c_double * 3)(*origin))
    nn, ne = C.c_int32(), C.c_int32()
    check(lib().cpi_mesh_cylinder(C.byref(s), C.byref(nn), C.byref(ne), None, None))
    npe = 10 if quadratic else 4
    x = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, npe), dtype=np.int32)
    check(lib().cpi_mesh_cylinder(C.byref(s), C.byref(nn), C.byref(ne), dptr(x), iptr(conn)))
    return x, conn


def greedy_coloring(conn, num_nodes):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    col = np.zeros(conn.shape[0], dtype=np.int32)
    n = check(lib().cpi_greedy_coloring(conn.shape[0], conn.shape[1], iptr(conn), num_nodes,
                                        iptr(col)))
    return col, n
