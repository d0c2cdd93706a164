"""ctypes mirror of include/vps.h (the C ABI of libvps).

The struct layouts here must match the header field for field; the library
exports ``vps_sizeof_params`` / ``vps_sizeof_stats`` and the binding checks
them when it loads.
"""
import ctypes as C
import os

VPS_OK = 0
VPS_ERR_INVALID_ARG = 1
VPS_ERR_NO_CONVERGENCE = 2
VPS_ERR_NONFINITE = 3
VPS_ERR_OOM = 4
VPS_ERR_HIP = 5
VPS_ERR_STATE = 6
VPS_ERR_IO = 7

STATUS_NAMES = {
    0: "VPS_OK",
    1: "VPS_ERR_INVALID_ARG",
    2: "VPS_ERR_NO_CONVERGENCE",
    3: "VPS_ERR_NONFINITE",
    4: "VPS_ERR_OOM",
    5: "VPS_ERR_HIP",
    6: "VPS_ERR_STATE",
    7: "VPS_ERR_IO",
}

OP_TRANSPOSE = 0
OP_ROW = 1

WARM_INDEX_CARRY = 0
WARM_ZERO = 1
WARM_NODE_KEYED = 2
WARM_LEARNED = 3
WARM_EXTRAPOLATE = 4

FROM_INIT = 0
FROM_PARAMS = 1

ABI_VERSION = 1


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("op_mode", C.c_int32),
        ("warm_start", C.c_int32),
        ("rest_density_mode", C.c_int32),
        ("kappa_mode", C.c_int32),
        ("reserved0", C.c_int32),
        ("particle_radius", C.c_double),
        ("lattice_spacing", C.c_double),
        ("lattice_origin", C.c_double * 3),
        ("lattice_dims", C.c_int32 * 3),
        ("reserved1", C.c_int32),
        ("mass", C.c_double),
        ("rest_density", C.c_double),
        ("kappa", C.c_double),
        ("tolerance", C.c_double),
        ("max_iterations", C.c_int64),
        ("domain_lo", C.c_double * 3),
        ("domain_hi", C.c_double * 3),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_real", C.c_int64),
        ("n_boundary", C.c_int64),
        ("n_virtual", C.c_int64),
        ("n_active_bricks", C.c_int64),
        ("cg_iterations", C.c_int64),
        ("cg_rel_residual", C.c_double),
        ("cg_true_rel_residual", C.c_double),
        ("rhs_norm", C.c_double),
        ("alpha0", C.c_double),
        ("kappa", C.c_double),
        ("rest_density", C.c_double),
        ("ms_sort", C.c_double),
        ("ms_generate", C.c_double),
        ("ms_diag", C.c_double),
        ("ms_cg", C.c_double),
        ("ms_gradient", C.c_double),
        ("ms_total", C.c_double),
        ("ms_spmv_avg", C.c_double),
        ("ms_update_avg", C.c_double),
        ("spmv_launches", C.c_int64),
        ("reserved", C.c_int64 * 4),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            if name == "reserved":
                continue
            out[name] = getattr(self, name)
        # residual replacements of the last solve, host time spent launching
        # its CG kernels (microseconds)
        out["cg_replacements"] = self.reserved[0]
        out["cg_host_launch_us"] = self.reserved[1]
        # 1 when the warm start overflowed and the solve ran again from the
        # plain (index-carried) start
        out["warm_start_fallback"] = self.reserved[2]
        return out


def make_params(scene, **overrides):
    """Builds a Params block from a scene dict (see scenes.py)."""
    p = Params()
    p.dim = scene["dim"]
    p.op_mode = overrides.pop("op_mode", OP_TRANSPOSE)
    p.warm_start = overrides.pop("warm_start", WARM_INDEX_CARRY)
    p.rest_density_mode = overrides.pop("rest_density_mode", scene.get("rest_density_mode", FROM_PARAMS))
    p.kappa_mode = overrides.pop("kappa_mode", scene.get("kappa_mode", FROM_PARAMS))
    p.particle_radius = scene["particle_radius"]
    p.lattice_spacing = scene["lattice_spacing"]
    for a in range(3):
        p.lattice_origin[a] = scene["lattice_origin"][a]
        p.lattice_dims[a] = int(scene["lattice_dims"][a])
        p.domain_lo[a] = scene["domain_lo"][a]
        p.domain_hi[a] = scene["domain_hi"][a]
    p.mass = scene["mass"]
    p.rest_density = scene["rest_density"]
    p.kappa = scene["kappa"]
    p.tolerance = overrides.pop("tolerance", 1e-6)
    p.max_iterations = overrides.pop("max_iterations", -1)
    if overrides:
        raise TypeError("unknown params: %s" % sorted(overrides))
    return p


def _proto(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I64 = C.POINTER(C.c_int64)


SURFACE_BOX, SURFACE_SPHERE, SURFACE_PLANE, SURFACE_ELLIPSOID = 0, 1, 2, 3


class Surface(C.Structure):
    """vps_surface: kind, complement, a[3], b[3] (include/vps.h)."""
    _fields_ = [("kind", C.c_int32), ("complement", C.c_int32),
                ("a", C.c_double * 3), ("b", C.c_double * 3)]


def surface(kind, a, b, complement=False):
    s = Surface()
    s.kind = kind
    s.complement = 1 if complement else 0
    for i in range(3):
        s.a[i] = float(a[i]) if i < len(a) else 0.0
        s.b[i] = float(b[i]) if i < len(b) else 0.0
    return s


class Collider(C.Structure):
    """vps_collider: a static collider's surface, restitution, friction."""
    _fields_ = [("surface", Surface), ("restitution", C.c_double),
                ("friction", C.c_double)]


def collider(kind, a, b, complement=False, restitution=0.0, friction=0.0):
    return Collider(surface(kind, a, b, complement), restitution, friction)


def collider_array(colliders):
    colliders = list(colliders)
    arr = (Collider * max(len(colliders), 1))()
    for i, c in enumerate(colliders):
        arr[i] = c
    return arr, len(colliders)


class StepParams(C.Structure):
    """vps_step_params (DualParticleSphLiquid::advance)."""
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double), ("viscosity", C.c_double),
                ("colliders", C.POINTER(Collider)), ("n_colliders", C.c_int32),
                ("enable_gravity", C.c_int32)]


class DemParams(C.Structure):
    """Contact constants of DEMParticle<Dim> (Cores/Structures/DEMParticle.cpp:
    15-60): Young's modulus (K_norm = young r), Poisson ratio
    (K_tang = K_norm poisson) and the friction angle in radians."""
    _fields_ = [
        ("young", C.c_double),
        ("poisson", C.c_double),
        ("friction_angle", C.c_double),
        ("reserved", C.c_double),
    ]


def dem_params(young=1e9, poisson=0.3, friction_angle=0.5):
    """The reference's defaults (DEMParticle.cpp:18-21)."""
    return DemParams(young, poisson, friction_angle, 0.0)


def declare(lib):
    """Declares every exported symbol of libvps on a CDLL handle."""
    _proto(lib, "vps_abi_version", C.c_int32, [])
    _proto(lib, "vps_sizeof_params", C.c_int64, [])
    _proto(lib, "vps_sizeof_stats", C.c_int64, [])
    _proto(lib, "vps_create", C.c_int, [C.POINTER(Params), C.POINTER(_P)])
    _proto(lib, "vps_destroy", None, [_P])
    _proto(lib, "vps_last_error", C.c_char_p, [_P])
    _proto(lib, "vps_set_boundary", C.c_int, [_P, C.c_int64, _D, _D, _D])
    _proto(lib, "vps_refresh_boundary", C.c_int, [_P])
    _proto(lib, "vps_set_particles", C.c_int, [_P, C.c_int64, _D, _D])
    _proto(lib, "vps_get_particles", C.c_int, [_P, _D, _D])
    _proto(lib, "vps_initialize", C.c_int, [_P, C.POINTER(Stats)])
    _proto(lib, "vps_refresh_real", C.c_int, [_P])
    _proto(lib, "vps_generate_particles", C.c_int, [_P])
    _proto(lib, "vps_project", C.c_int, [_P, C.POINTER(Stats)])
    _proto(lib, "vps_projection_step", C.c_int, [_P, C.POINTER(Stats)])
    _proto(lib, "vps_advect", C.c_int, [_P, C.c_double, C.c_double])
    _proto(lib, "vps_set_check_true_residual", C.c_int, [_P, C.c_int32])
    _proto(lib, "vps_set_async_steps", C.c_int, [_P, C.c_int32])
    _proto(lib, "vps_synchronize", C.c_int, [_P])
    _proto(lib, "vps_get_virtual_count", C.c_int, [_P, _I64])
    _proto(lib, "vps_get_virtual", C.c_int, [_P, _I64, _D, _D, _D, _D, _D])
    _proto(lib, "vps_get_real_fields", C.c_int, [_P, _D, _D])
    _proto(lib, "vps_get_boundary_fields", C.c_int, [_P, _D, _D])
    _proto(lib, "vps_get_scalars", C.c_int, [_P, _D, _D, _D])
    _proto(lib, "vps_query_neighbors", C.c_int,
           [_P, C.c_int32, C.c_int64, _D, _I64, _I64, C.c_int64, _I64])
    _proto(lib, "vps_mlp_set_training", C.c_int,
           [_P, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_uint64])
    _proto(lib, "vps_mlp_parameter_count", C.c_int64, [])
    _proto(lib, "vps_mlp_get_parameters", C.c_int, [_P, _F])
    _proto(lib, "vps_mlp_set_parameters", C.c_int, [_P, _F])
    _proto(lib, "vps_mlp_loss_and_grad", C.c_int, [_P, C.c_int64, _F, _F, _D, _F])
    _proto(lib, "vps_mlp_loss_and_grad_device", C.c_int,
           [_P, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p])
    _proto(lib, "vps_mlp_forward", C.c_int, [_P, C.c_int64, _F, _F])
    _proto(lib, "vps_mlp_get_features", C.c_int, [_P, _F])
    _proto(lib, "vps_mlp_training_state", C.c_int, [_P, _D, _I64])
    _proto(lib, "vps_dem_step", C.c_int,
           [_P, C.c_double, C.c_double, C.POINTER(DemParams)])
    _proto(lib, "vps_couple_grains", C.c_int, [_P, _P])
    _proto(lib, "vps_move_particles", C.c_int, [_P, C.c_double, C.POINTER(Collider), C.c_int32])
    _proto(lib, "vps_apply_gravity", C.c_int, [_P, C.c_double, C.c_double])
    _proto(lib, "vps_apply_viscosity", C.c_int, [_P, C.c_double, C.c_double])
    _proto(lib, "vps_advance", C.c_int, [_P, C.POINTER(StepParams), C.POINTER(Stats)])
    _proto(lib, "vps_add_surface", C.c_int, [_P, C.POINTER(Surface), _D, _D])
    _proto(lib, "vps_generate_box", C.c_int, [_P, _D, _D, C.c_int64, _D, _I64])
    _proto(lib, "vps_get_boundary", C.c_int, [_P, _I64, _D, _D, _D])
    _proto(lib, "vps_write_frame", C.c_int, [_P, C.c_char_p])
    _proto(lib, "vps_save_frame", C.c_int, [_P, C.c_char_p])
    _proto(lib, "vps_load_frame", C.c_int, [_P, C.c_char_p])
    _proto(lib, "vps_write_description", C.c_int, [C.c_int32, C.c_char_p])
    _proto(lib, "vps_write_end_frame", C.c_int, [C.c_char_p, C.c_int32])
    _proto(lib, "vps_write_and_save_frame", C.c_int, [_P, C.c_char_p, C.c_int32])
    _proto(lib, "vps_load_frame_number", C.c_int, [_P, C.c_char_p, C.c_int32])
    _proto(lib, "vps_debug_apply_operator", C.c_int, [_P, _D, _D])
    _proto(lib, "vps_debug_member_counts", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)])
    return lib


# Every symbol include/vps.h declares (checked by the CPU test suite).
EXPORTED = [
    "vps_abi_version", "vps_sizeof_params", "vps_sizeof_stats", "vps_create",
    "vps_destroy", "vps_last_error", "vps_set_boundary", "vps_refresh_boundary",
    "vps_set_particles", "vps_get_particles", "vps_initialize",
    "vps_refresh_real", "vps_generate_particles", "vps_project",
    "vps_projection_step", "vps_advect", "vps_set_check_true_residual",
    "vps_set_async_steps",
    "vps_synchronize", "vps_get_virtual_count", "vps_get_virtual",
    "vps_get_real_fields", "vps_get_boundary_fields", "vps_get_scalars",
    "vps_query_neighbors", "vps_mlp_set_training", "vps_mlp_parameter_count",
    "vps_mlp_get_parameters", "vps_mlp_set_parameters",
    "vps_mlp_loss_and_grad", "vps_mlp_loss_and_grad_device", "vps_mlp_forward",
    "vps_mlp_get_features",
    "vps_mlp_training_state", "vps_dem_step", "vps_couple_grains",
    "vps_add_surface", "vps_generate_box", "vps_get_boundary",
    "vps_move_particles", "vps_apply_gravity", "vps_apply_viscosity", "vps_advance",
    "vps_write_frame", "vps_save_frame", "vps_load_frame", "vps_write_description",
    "vps_write_end_frame", "vps_write_and_save_frame", "vps_load_frame_number",
    "vps_debug_apply_operator", "vps_debug_member_counts",
]

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libvps.so")
