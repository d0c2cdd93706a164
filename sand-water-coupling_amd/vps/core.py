"""ctypes wrapper around libvps (see include/vps.h)."""
import ctypes as C
import os

import numpy as np

from . import _abi

_lib = None


class VpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_abi.STATUS_NAMES.get(code, code), msg))
        self.code = code


def load_library(path=None):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or _abi.LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            "libvps.so not found at %s: build it with __graft_entry__.build() "
            "(make -C sand-water-coupling_amd/csrc)" % path)
    lib = C.CDLL(path)
    _abi.declare(lib)
    lib.vps_set_kernel_timing.restype = C.c_int
    lib.vps_set_kernel_timing.argtypes = [C.c_void_p, C.c_int32]
    if lib.vps_sizeof_params() != C.sizeof(_abi.Params):
        raise RuntimeError("vps_params layout mismatch")
    if lib.vps_sizeof_stats() != C.sizeof(_abi.Stats):
        raise RuntimeError("vps_stats layout mismatch")
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Projection:
    """One projection context on the current GPU.

    Call order mirrors the reference: ``initialize()`` once
    (DualParticleSphLiquid::initialize), then per step ``refresh_real()``,
    ``generate_particles()`` and ``project()`` -- or ``projection_step()``
    for all three.
    """

    def __init__(self, scene, **overrides):
        self.L = load_library()
        self.dim = scene["dim"]
        self.scene = scene
        self.params = _abi.make_params(scene, **overrides)
        h = C.c_void_p()
        rc = self.L.vps_create(C.byref(self.params), C.byref(h))
        if rc:
            raise VpsError(rc, self.L.vps_last_error(None).decode())
        self.h = h
        bx = np.ascontiguousarray(scene["bx"], dtype=np.float64)
        bn = np.ascontiguousarray(scene["bn"], dtype=np.float64)
        bv = np.ascontiguousarray(scene["bv"], dtype=np.float64)
        self._check(self.L.vps_set_boundary(self.h, len(bx), _d(bx), _d(bn), _d(bv)))
        self.nb = len(bx)
        self.set_particles(scene["x"], scene["v"])

    # -- plumbing
    def _check(self, rc):
        if rc:
            raise VpsError(rc, self.L.vps_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.vps_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state
    def set_particles(self, x, v=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.n = len(x)
        if v is None:
            self._check(self.L.vps_set_particles(self.h, self.n, _d(x), None))
        else:
            v = np.ascontiguousarray(v, dtype=np.float64)
            self._check(self.L.vps_set_particles(self.h, self.n, _d(x), _d(v)))

    def particles(self):
        x = np.zeros((self.n, self.dim))
        v = np.zeros((self.n, self.dim))
        self._check(self.L.vps_get_particles(self.h, _d(x), _d(v)))
        return x, v

    # -- path
    def initialize(self):
        st = _abi.Stats()
        self._check(self.L.vps_initialize(self.h, C.byref(st)))
        return st.as_dict()

    def refresh_real(self):
        self._check(self.L.vps_refresh_real(self.h))

    def generate_particles(self):
        self._check(self.L.vps_generate_particles(self.h))

    def project(self, allow_fail=False):
        st = _abi.Stats()
        rc = self.L.vps_project(self.h, C.byref(st))
        if not allow_fail:
            self._check(rc)
        d = st.as_dict()
        d["status"] = rc
        return d

    def projection_step(self):
        st = _abi.Stats()
        self._check(self.L.vps_projection_step(self.h, C.byref(st)))
        return st.as_dict()

    def advect(self, dt, gravity):
        self._check(self.L.vps_advect(self.h, dt, gravity))

    def synchronize(self):
        self._check(self.L.vps_synchronize(self.h))

    def set_check_true_residual(self, on):
        self._check(self.L.vps_set_check_true_residual(self.h, 1 if on else 0))

    def set_async_steps(self, on):
        """Steps return once queued (no wait for their end; zero phase times)."""
        self._check(self.L.vps_set_async_steps(self.h, 1 if on else 0))

    def set_kernel_timing(self, on):
        self._check(self.L.vps_set_kernel_timing(self.h, 1 if on else 0))

    # -- read-back
    def virtual_count(self):
        n = C.c_int64()
        self._check(self.L.vps_get_virtual_count(self.h, C.byref(n)))
        return n.value

    def virtual(self):
        n = self.virtual_count()
        out = {k: np.zeros(n) for k in ("rho", "vol", "div", "diag", "p")}
        out["node"] = np.zeros(n, dtype=np.int64)
        self._check(self.L.vps_get_virtual(
            self.h, _i64(out["node"]), _d(out["rho"]), _d(out["vol"]),
            _d(out["div"]), _d(out["diag"]), _d(out["p"])))
        return out

    def real_fields(self):
        rho = np.zeros(self.n)
        vol = np.zeros(self.n)
        self._check(self.L.vps_get_real_fields(self.h, _d(rho), _d(vol)))
        return rho, vol

    def boundary_fields(self):
        rho = np.zeros(self.nb)
        vol = np.zeros(self.nb)
        self._check(self.L.vps_get_boundary_fields(self.h, _d(rho), _d(vol)))
        return rho, vol

    def scalars(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.vps_get_scalars(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"rest_density": a.value, "kappa": b.value, "alpha0": c.value}

    def query(self, which, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, self.dim)
        nq = len(q)
        offs = np.zeros(nq + 1, dtype=np.int64)
        need = C.c_int64(0)
        self._check(self.L.vps_query_neighbors(self.h, which, nq, _d(q), _i64(offs),
                                               None, 0, C.byref(need)))
        idx = np.zeros(max(need.value, 1), dtype=np.int64)
        self._check(self.L.vps_query_neighbors(self.h, which, nq, _d(q), _i64(offs),
                                               _i64(idx), len(idx), C.byref(need)))
        return offs, idx[: need.value]

    def member_counts(self, which):
        """Terms per particle of the FP32 walks (0 real density, 1 boundary
        density, 2 the gradient's virtual walk, 3 viscosity), caller order."""
        n = self.nb if which == 1 else self.n
        out = np.zeros(n, dtype=np.int32)
        self._check(self.L.vps_debug_member_counts(
            self.h, int(which), out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def apply_operator(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.zeros_like(p)
        self._check(self.L.vps_debug_apply_operator(self.h, _d(p), _d(out)))
        return out

    # -- learned warm start (no reference counterpart)
    def mlp_set_training(self, enable=True, batch=65536, lr=1e-3,
                         steps_per_solve=1, seed=1):
        self._check(self.L.vps_mlp_set_training(self.h, 1 if enable else 0,
                                                int(batch), float(lr),
                                                int(steps_per_solve), int(seed)))

    def mlp_parameters(self):
        p = np.zeros(self.L.vps_mlp_parameter_count(), dtype=np.float32)
        self._check(self.L.vps_mlp_get_parameters(self.h, _f(p)))
        return p

    def mlp_set_parameters(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.L.vps_mlp_parameter_count():
            raise ValueError("expected %d parameters" % self.L.vps_mlp_parameter_count())
        self._check(self.L.vps_mlp_set_parameters(self.h, _f(p)))

    def mlp_loss_and_grad(self, features, targets):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        if len(t) != len(f):
            raise ValueError("features and targets differ in length")
        loss = C.c_double()
        g = np.zeros(self.L.vps_mlp_parameter_count(), dtype=np.float32)
        self._check(self.L.vps_mlp_loss_and_grad(self.h, len(f), _f(f), _f(t),
                                                 C.byref(loss), _f(g)))
        return loss.value, g

    def mlp_loss_and_grad_device(self, n, d_features, d_targets, d_loss, d_grad):
        """vps_mlp_loss_and_grad_device on device pointers (integers): queues
        the kernels on the context's stream and returns; call synchronize()
        before reading d_loss / d_grad."""
        self._check(self.L.vps_mlp_loss_and_grad_device(self.h, int(n), d_features,
                                                        d_targets, d_loss, d_grad))

    def mlp_forward(self, features):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(len(f), dtype=np.float32)
        self._check(self.L.vps_mlp_forward(self.h, len(f), _f(f), _f(out)))
        return out

    def mlp_features(self):
        out = np.zeros((self.virtual_count(), 8), dtype=np.float32)
        self._check(self.L.vps_mlp_get_features(self.h, _f(out)))
        return out

    def mlp_training_state(self):
        loss, steps = C.c_double(), C.c_int64()
        self._check(self.L.vps_mlp_training_state(self.h, C.byref(loss), C.byref(steps)))
        return loss.value, steps.value

    # -- the rest of the SPH step (SmthParticleHydrodLiquid / DualParticleSphLiquid)
    def move_particles(self, dt, colliders=()):
        arr, n = _abi.collider_array(colliders)
        self._check(self.L.vps_move_particles(self.h, dt, arr, n))

    def apply_gravity(self, dt, gravity=9.8):
        self._check(self.L.vps_apply_gravity(self.h, dt, gravity))

    def apply_viscosity(self, dt, coefficient=0.01):
        self._check(self.L.vps_apply_viscosity(self.h, dt, coefficient))

    def advance(self, dt, gravity=9.8, viscosity=0.01, colliders=(), enable_gravity=False):
        """DualParticleSphLiquid::advance: move (+ colliders), gravity when
        enabled, viscosity, generation, projection."""
        arr, n = _abi.collider_array(colliders)
        sp = _abi.StepParams(dt, gravity, viscosity, arr, n, 1 if enable_gravity else 0)
        st = _abi.Stats()
        self._check(self.L.vps_advance(self.h, C.byref(sp), C.byref(st)))
        return st.as_dict()

    # -- wall and block sampling (BoundaryParticles::addSurface, Shapes::generateBox)
    def add_surface(self, kind, a, b, center, half_lengths, complement=False):
        s = _abi.surface(kind, a, b, complement)
        c = np.ascontiguousarray(center, dtype=np.float64)
        h = np.ascontiguousarray(half_lengths, dtype=np.float64)
        self._check(self.L.vps_add_surface(self.h, C.byref(s), _d(c), _d(h)))
        self.nb = self.boundary()[0].shape[0]

    def generate_box(self, center, half_lengths):
        c = np.ascontiguousarray(center, dtype=np.float64)
        h = np.ascontiguousarray(half_lengths, dtype=np.float64)
        cnt = C.c_int64()
        self._check(self.L.vps_generate_box(self.h, _d(c), _d(h), 0, None, C.byref(cnt)))
        x = np.zeros((cnt.value, self.dim))
        self._check(self.L.vps_generate_box(self.h, _d(c), _d(h), cnt.value, _d(x),
                                            C.byref(cnt)))
        return x

    def boundary(self):
        """The walls as given: positions, normals, velocities."""
        cnt = C.c_int64()
        self._check(self.L.vps_get_boundary(self.h, C.byref(cnt), None, None, None))
        x = np.zeros((cnt.value, self.dim))
        nrm = np.zeros((cnt.value, self.dim))
        v = np.zeros((cnt.value, self.dim))
        self._check(self.L.vps_get_boundary(self.h, C.byref(cnt), _d(x), _d(nrm), _d(v)))
        return x, nrm, v

    # -- frames and checkpoints (reference file formats, see vps.frames)
    def write_frame(self, frame_dir):
        self._check(self.L.vps_write_frame(self.h, os.fsencode(frame_dir)))

    def save_frame(self, frame_dir):
        self._check(self.L.vps_save_frame(self.h, os.fsencode(frame_dir)))

    def load_frame(self, frame_dir):
        self._check(self.L.vps_load_frame(self.h, os.fsencode(frame_dir)))
        self.n = os.path.getsize(os.path.join(frame_dir, "particles.sav")) // (8 * self.dim)

    def write_and_save_frame(self, output_dir, frame):
        """Simulator::writeAndSaveToFrameDirectory: <output_dir>/<frame>/."""
        self._check(self.L.vps_write_and_save_frame(self.h, os.fsencode(output_dir), frame))

    def load_frame_number(self, output_dir, frame):
        self._check(self.L.vps_load_frame_number(self.h, os.fsencode(output_dir), frame))
        self.n = os.path.getsize(os.path.join(output_dir, str(frame), "particles.sav")) // (
            8 * self.dim)

    # -- grains as moving boundary particles
    def couple_grains(self, grains):
        """Boundary = this context's walls followed by the grains of a
        ``DemSand`` context (zero normals: the pressure-gradient boundary term
        takes the unit vector from the grain to the fluid particle)."""
        self._check(self.L.vps_couple_grains(self.h, grains.h))
        self.nb = len(self.scene["bx"]) + grains.n


class DemSand(Projection):
    """The grains of DEMParticleSand<Dim> (Cores/Physics/DEMParticleSand.h:13-58)
    on the GPU: the scene's particles are the grains, its boundary particles
    the walls.  ``advance(dt)`` is DEMParticleSand::advance
    (DEMParticleSand.cpp:90-94): move, contact and wall forces, gravity."""

    def __init__(self, scene, young=1e9, poisson=0.3, friction_angle=0.5,
                 gravity=9.8, **overrides):
        super().__init__(scene, **overrides)
        self.dem = _abi.dem_params(young, poisson, friction_angle)
        self.gravity = gravity

    def advance(self, dt):
        self._check(self.L.vps_dem_step(self.h, dt, self.gravity, C.byref(self.dem)))


def write_description(dim, output_dir):
    """description.yaml of an output directory (Simulator.cpp:33-39)."""
    L = load_library()
    rc = L.vps_write_description(dim, os.fsencode(output_dir))
    if rc:
        raise VpsError(rc, "cannot write %s/description.yaml" % output_dir)
