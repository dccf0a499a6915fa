"""Host-side mirror of the reference solver interface over the C ABI.

The reference's scene + step API is the C++ class pair SPHSolver /
PCISPHSolver (LiquidPhysics/include/SPHSolver.h:25-69, PCISPHSolver.h:7-20):
particles and static colliders are added, parameters are set through setters,
and PhysicSolver::update(dt) advances one step (PhysicSolver.cpp:12-21).  The
classes below keep those method names and meanings, with particles held as
arrays (the library keeps the state on the device, in the reference's sorted
order, like its OpenCL path; SPHSolver.cpp:412-422) instead of SPHParticle
objects.  Every call goes through liblps.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional

import numpy as np

from . import _lib
from .abi import (CBox, CGridInfo, CParams, CPlane, CSphere, CStepStats, CTrainCfg, CTuning, Box, Colliders, Params, Plane,
                  Sphere, LPS_TABLE_SIZE, LPS_MLP_IN, LPS_MLP_PARAMS, METHOD_PCISPH, METHOD_SPH)
from ._lib import check, lib, ptr


def device_info():
    name = C.create_string_buffer(256)
    cus = C.c_int32()
    st = lib().lps_device_info(name, 256, C.byref(cus))
    check(st, "lps_device_info")
    return name.value.decode(), cus.value


class SPHSolver:
    """SPH solver backed by the gfx950 library (reference: SPHSolver)."""

    _default_method = METHOD_SPH

    def __init__(self, params: Optional[Params] = None):
        self._params = params if params is not None else Params(method=self._default_method)
        self._ctx = C.c_void_p()
        cp = self._params.to_c()
        check(lib().lps_create(C.byref(cp), C.byref(self._ctx)), "lps_create")
        self._colliders = Colliders()
        self._pending: List[np.ndarray] = []
        self._n = 0
        self.last_stats: List[CStepStats] = []

    # ---- lifetime
    def close(self):
        if self._ctx:
            lib().lps_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters (SPHSolver setters, SPHSolver.cpp:628-710)
    @property
    def params(self) -> Params:
        return self._params

    def _apply(self, **kw):
        self.set_params(self._params.replace(**kw))

    def set_params(self, params: Params):
        # kept only once the library accepted them
        cp = params.to_c()
        check(lib().lps_set_params(self._ctx, C.byref(cp)), "lps_set_params")
        self._params = params

    def setGravity(self, g):
        self._apply(gravity=tuple(float(v) for v in g))

    def setParticleRadius(self, r):
        self._apply(particle_radius=float(r))

    def setKernelRadiusFactor(self, f):
        self._apply(kernel_radius_factor=float(f))

    def setRestDensity(self, v):
        self._apply(rest_density=float(v))

    def setViscosityCoefficient(self, v):
        self._apply(viscosity=float(v))

    def setPressureStiffnessCoefficient(self, v):
        self._apply(stiffness=float(v))

    def setNegativePressureFactor(self, v):
        self._apply(negative_pressure_factor=float(v))

    def setSurfaceTensionCoefficient(self, v):
        self._apply(surface_tension=float(v))

    def setSurfaceTensionThreshold(self, v):
        self._apply(surface_tension_threshold=float(v))

    def setRestitutionCoefficient(self, v):
        self._apply(restitution=float(v))

    def setFrictionCoefficient(self, v):
        self._apply(friction=float(v))

    # ---- getters
    def getParticleCount(self) -> int:
        return self._n

    def getParticleRadius(self):
        return self._params.particle_radius

    def _derived(self):
        h, ds, m = C.c_float(), C.c_float(), C.c_float()
        t = np.zeros(5 * LPS_TABLE_SIZE, np.float32)
        check(lib().lps_get_derived(self._ctx, C.byref(h), C.byref(ds), C.byref(m), t.ctypes.data),
              "lps_get_derived")
        return h.value, ds.value, m.value, t.reshape(5, LPS_TABLE_SIZE)

    def getKernelRadius(self):
        return self._derived()[0]

    def getParticleMass(self):
        return self._derived()[2]

    def kernel_tables(self):
        return self._derived()[3]

    def division_step(self):
        return self._derived()[1]

    # ---- scene (addParticle / addStaticCollisionObject / cleanUp)
    def set_particles(self, x, v=None, vhalf=None, first=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        n = x.shape[0]
        v = None if v is None else np.ascontiguousarray(v, np.float32).reshape(n, 3)
        vhalf = None if vhalf is None else np.ascontiguousarray(vhalf, np.float32).reshape(n, 3)
        first = None if first is None else np.ascontiguousarray(first, np.uint8).reshape(n)
        self._push_colliders()
        check(lib().lps_set_particles(self._ctx, x.reshape(-1), ptr(v), ptr(vhalf), ptr(first), n),
              "lps_set_particles")
        self._n = n
        self._pending = []

    def add_particles(self, x, v=None, vhalf=None, first=None):
        """Append particles to the running state on the device (lps_add_particles);
        the current particles keep their state and sorted order."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        n = x.shape[0]
        if n == 0:
            return
        if not self._n:
            self.set_particles(x, v, vhalf, first)
            return
        va = None if v is None else np.ascontiguousarray(v, np.float32).reshape(n, 3)
        ha = None if vhalf is None else np.ascontiguousarray(vhalf, np.float32).reshape(n, 3)
        fa = None if first is None else np.ascontiguousarray(first, np.uint8).reshape(n)
        self._push_colliders()
        check(lib().lps_add_particles(self._ctx, x.reshape(-1), ptr(va), ptr(ha), ptr(fa), n), "lps_add_particles")
        self._n += n

    def positions_into(self, dst) -> None:
        """Writes the particle positions (get_particles order) into a CUDA
        tensor of shape [n, k >= 3], float32, contiguous rows; no host copy
        (lps_copy_positions_device)."""
        import torch
        self._flush_pending()
        if not (isinstance(dst, torch.Tensor) and dst.is_cuda and dst.dtype == torch.float32 and dst.dim() == 2
                and dst.shape[0] == self._n and dst.shape[1] >= 3 and dst.stride(1) == 1):
            raise ValueError("positions_into needs a CUDA float32 tensor [n, k >= 3] with contiguous rows")
        check(lib().lps_copy_positions_device(self._ctx, dst.data_ptr(), dst.stride(0), self._n,
                                              dst.untyped_storage().nbytes() - dst.storage_offset() * 4),
              "lps_copy_positions_device")

    def addParticles(self, x, v=None):
        """Queue particles (positions, velocities); uploaded at the next update."""
        x = np.asarray(x, np.float32).reshape(-1, 3)
        v = np.zeros_like(x) if v is None else np.asarray(v, np.float32).reshape(-1, 3)
        self._pending.append((x, v))

    def removeParticles(self):
        self._pending = []
        self._n = 0

    def addStaticCollisionObject(self, obj):
        if isinstance(obj, Box):
            self._colliders.boxes.append(obj)
        elif isinstance(obj, Sphere):
            self._colliders.spheres.append(obj)
        elif isinstance(obj, Plane):
            self._colliders.planes.append(obj)
        else:
            raise TypeError(type(obj))
        self._colliders_dirty = True

    def set_colliders(self, col: Colliders):
        self._colliders = col
        self._colliders_dirty = True
        self._push_colliders()

    def removeStaticCollisionObjects(self):
        self._colliders = Colliders()
        self._colliders_dirty = True

    def cleanUp(self):
        self.removeParticles()
        self.removeStaticCollisionObjects()

    def _push_colliders(self):
        if not getattr(self, "_colliders_dirty", True):
            return
        boxes, nb, spheres, ns, planes, npl = self._colliders.c_arrays()
        check(lib().lps_set_colliders(self._ctx, C.cast(boxes, C.c_void_p), nb, C.cast(spheres, C.c_void_p), ns,
                                      C.cast(planes, C.c_void_p), npl), "lps_set_colliders")
        self._colliders_dirty = False

    def _flush_pending(self):
        if self._pending:
            # queued particles follow the current ones, as SPHSolver::addParticle
            # appends to the particle list
            x = np.concatenate([p[0] for p in self._pending])
            v = np.concatenate([p[1] for p in self._pending])
            self._pending = []
            self.add_particles(x, v)
        self._push_colliders()

    # ---- step (PhysicSolver::update)
    def update(self, dt: float, nsteps: int = 1):
        self._flush_pending()
        stats = (CStepStats * max(nsteps, 1))()
        st = lib().lps_step(self._ctx, float(dt), int(nsteps), C.cast(stats, C.c_void_p))
        self.last_stats = list(stats[:nsteps])
        check(st, "lps_step")
        return self.last_stats

    # ---- readback (onEndUpdate) and parity hooks
    def get_particles(self):
        n = self._n
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        vh = np.zeros((n, 3), np.float32)
        first = np.zeros(n, np.uint8)
        ids = np.zeros(n, np.uint32)
        check(lib().lps_get_particles(self._ctx, ptr(x), ptr(v), ptr(vh), ptr(first), ptr(ids), n),
              "lps_get_particles")
        return {"x": x, "v": v, "vhalf": vh, "first": first, "id": ids}

    def get_grid(self, cell_cap: Optional[int] = None):
        info = CGridInfo()
        check(lib().lps_get_grid(self._ctx, C.byref(info), None, None, None, None, 0), "lps_get_grid")
        cells = info.cell_count if cell_cap is None else min(cell_cap, info.cell_count)
        n = self._n
        keys = np.zeros(n, np.uint32)
        perm = np.zeros(n, np.uint32)
        cs = np.zeros(max(cells, 1), np.uint32)
        ce = np.zeros(max(cells, 1), np.uint32)
        check(lib().lps_get_grid(self._ctx, C.byref(info), ptr(keys), ptr(perm), ptr(cs), ptr(ce), cells),
              "lps_get_grid")
        return {
            "bbox_min": np.array(info.bbox_min[:], np.float32),
            "bbox_max": np.array(info.bbox_max[:], np.float32),
            "grid_offset": np.array(info.grid_offset[:], np.float32),
            "grid_spacing": info.grid_spacing,
            "grid_size": np.array(info.grid_size[:], np.uint32),
            "cell_count": info.cell_count,
            "keys": keys, "perm": perm, "cell_start": cs[:cells], "cell_end": ce[:cells],
        }

    def get_fields(self):
        n = self._n
        rho = np.zeros(n, np.float32)
        p = np.zeros(n, np.float32)
        f = np.zeros((n, 3), np.float32)
        rp = np.zeros(n, np.float32)
        check(lib().lps_get_fields(self._ctx, ptr(rho), ptr(p), ptr(f), ptr(rp)), "lps_get_fields")
        return {"rho": rho, "pressure": p, "force": f, "rho_pred": rp}

    def get_energy(self):
        """(kinetic, potential) energy of the current state, per scene
        (float64 arrays), reduced on the device."""
        nsc = C.c_uint32()
        check(lib().lps_get_scene_count(self._ctx, C.byref(nsc)), "lps_get_scene_count")
        ke = np.zeros(nsc.value, np.float64)
        pe = np.zeros(nsc.value, np.float64)
        check(lib().lps_get_energy(self._ctx, ke.ctypes.data, pe.ctypes.data, nsc.value), "lps_get_energy")
        return ke, pe

    def delta(self, dt: float) -> float:
        d = C.c_float()
        check(lib().lps_get_delta(self._ctx, float(dt), C.byref(d)), "lps_get_delta")
        return d.value

    # ---- launch tuning
    def set_tuning(self, use_graphs: bool, lookahead: int = 1, correction_threads: int = 256,
                   lanes_per_particle: int = 0):
        t = CTuning(use_graphs=int(bool(use_graphs)), lookahead=int(lookahead), ms_per_step=0.0,
                    correction_threads=int(correction_threads), lanes_per_particle=int(lanes_per_particle))
        check(lib().lps_set_tuning(self._ctx, C.byref(t)), "lps_set_tuning")

    def get_tuning(self):
        t = CTuning()
        check(lib().lps_get_tuning(self._ctx, C.byref(t)), "lps_get_tuning")
        return {"use_graphs": bool(t.use_graphs), "lookahead": t.lookahead,
                "correction_threads": t.correction_threads, "lanes_per_particle": t.lanes_per_particle}

    def autotune(self, dt: float, steps: int = 10):
        """Time every launch setting for `steps` steps from the current state,
        restore the state and keep the fastest setting."""
        self._flush_pending()
        t = CTuning()
        check(lib().lps_autotune(self._ctx, float(dt), int(steps), C.byref(t)), "lps_autotune")
        return {"use_graphs": bool(t.use_graphs), "lookahead": t.lookahead, "ms_per_step": t.ms_per_step,
                "correction_threads": t.correction_threads, "lanes_per_particle": t.lanes_per_particle}

    # ---- timing
    def profile(self, on: bool = True):
        check(lib().lps_profile_enable(self._ctx, 1 if on else 0), "lps_profile_enable")

    def profile_reset(self):
        check(lib().lps_profile_reset(self._ctx), "lps_profile_reset")

    def profile_query(self, name: str):
        n = C.c_uint32()
        ms = C.c_double()
        check(lib().lps_profile_query(self._ctx, name.encode(), C.byref(n), C.byref(ms)), "lps_profile_query")
        return n.value, ms.value

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().lps_get_stream(self._ctx, C.byref(s)), "lps_get_stream")
        return s.value or 0


class PCISPHSolver(SPHSolver):
    """PCISPH solver (reference: PCISPHSolver, PCISPHSolver.cpp:50-240), plus
    the learned pressure warm start (no counterpart in the reference)."""

    _default_method = METHOD_PCISPH

    # ---- learned warm start
    def set_warm_start_weights(self, params, feature_scale, target_scale: float):
        params = np.ascontiguousarray(params, np.float32).reshape(LPS_MLP_PARAMS)
        feature_scale = np.ascontiguousarray(feature_scale, np.float32).reshape(2 * LPS_MLP_IN)
        check(lib().lps_mlp_set_weights(self._ctx, params, feature_scale, float(target_scale)),
              "lps_mlp_set_weights")

    def get_warm_start_weights(self):
        w = np.zeros(LPS_MLP_PARAMS, np.float32)
        fs = np.zeros(2 * LPS_MLP_IN, np.float32)
        ts = C.c_float()
        check(lib().lps_mlp_get_weights(self._ctx, w, fs, C.byref(ts)), "lps_mlp_get_weights")
        return {"params": w, "feature_scale": fs, "target_scale": ts.value}

    def capture_features(self, on: bool = True):
        check(lib().lps_mlp_capture(self._ctx, 1 if on else 0), "lps_mlp_capture")

    def dump_pairs(self, seed: bool = False):
        """Features [n, 16], applied pressure [n] (and the network's seed [n],
        before the envelope) of the last step."""
        n = self._n
        x = np.zeros((n, LPS_MLP_IN), np.float32)
        t = np.zeros(n, np.float32)
        s = np.zeros(n, np.float32) if seed else None
        check(lib().lps_dump_pairs(self._ctx, ptr(x), ptr(t), ptr(s), n), "lps_dump_pairs")
        return (x, t, s) if seed else (x, t)

    def collect(self, capacity: int, stride: int = 1):
        check(lib().lps_mlp_collect(self._ctx, int(capacity), int(stride)), "lps_mlp_collect")

    def collected(self) -> int:
        r = C.c_uint64()
        check(lib().lps_mlp_collected(self._ctx, C.byref(r)), "lps_mlp_collected")
        return r.value

    def train_warm_start(self, epochs: int, batch: int = 65536, lr: float = 2e-3, lr_final: float = 1e-5,
                         beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, reinit: bool = True,
                         refit_scale: bool = True, over_weight: float = 1.0, seed: int = 0):
        cfg = CTrainCfg(epochs=int(epochs), batch=int(batch), lr=lr, lr_final=lr_final, beta1=beta1, beta2=beta2,
                        eps=eps, reinit=int(reinit), refit_scale=int(refit_scale), over_weight=float(over_weight),
                        seed=int(seed))
        loss = np.zeros(max(int(epochs), 1), np.float32)
        check(lib().lps_mlp_train(self._ctx, C.byref(cfg), loss.ctypes.data), "lps_mlp_train")
        return loss[: int(epochs)]

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, LPS_MLP_IN)
        y = np.zeros(x.shape[0], np.float32)
        check(lib().lps_mlp_forward(self._ctx, x.reshape(-1), x.shape[0], y), "lps_mlp_forward")
        return y

    def mlp_grad(self, x, t, over_weight: float = 1.0):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, LPS_MLP_IN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        g = np.zeros(LPS_MLP_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().lps_mlp_grad(self._ctx, x.reshape(-1), t, x.shape[0], float(over_weight), g, C.byref(loss)),
              "lps_mlp_grad")
        return g, loss.value

    def mlp_grad_noalloc(self, x, t, over_weight: float = 1.0):
        """mlp_grad through lps_mlp_grad_noalloc (no device allocation; at most
        LPS_NOALLOC_MAX_ROWS rows)."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, LPS_MLP_IN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        g = np.zeros(LPS_MLP_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().lps_mlp_grad_noalloc(self._ctx, x.reshape(-1), t, x.shape[0], float(over_weight), g,
                                         C.byref(loss)), "lps_mlp_grad_noalloc")
        return g, loss.value

    def setMinIterations(self, k: int):
        self._apply(min_iterations=int(k))

    def setMaxIterations(self, k: int):
        self._apply(max_iterations=int(k))

    def setMaxDensityErrorRatio(self, r: float):
        self._apply(max_density_error_ratio=float(r))

    def getMinIterations(self):
        return self._params.min_iterations

    def getMaxDensityErrorRatio(self):
        return self._params.max_density_error_ratio


class Ensemble(PCISPHSolver):
    """Many independent scenes in one context (no counterpart in the
    reference): shared fluid parameters, per-scene gravity and colliders, one
    launch per phase for all scenes.  Each scene's results equal those of the
    scene run alone.  Works for SPH and PCISPH (params.method)."""

    def set_scenes(self, xs, vs=None, gravity=None):
        """xs: list of [n_s, 3] position arrays; vs: matching velocities or None;
        gravity: [S, 3] or None (the parameters' gravity for every scene)."""
        S = len(xs)
        counts = np.array([np.asarray(x).reshape(-1, 3).shape[0] for x in xs], np.uint32)
        x = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1, 3) for a in xs]))
        v = None
        if vs is not None:
            v = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1, 3) for a in vs]))
        g = None if gravity is None else np.ascontiguousarray(np.asarray(gravity, np.float32).reshape(S, 3))
        self._push_colliders()
        check(lib().lps_set_scenes(self._ctx, S, counts, ptr(g), x.reshape(-1), ptr(v), None, None),
              "lps_set_scenes")
        self._n = int(counts.sum())
        self._counts = counts
        self._pending = []

    @property
    def scene_count(self) -> int:
        n = C.c_uint32()
        check(lib().lps_get_scene_count(self._ctx, C.byref(n)), "lps_get_scene_count")
        return n.value

    def scene_slices(self):
        """Slices of each scene's particles in the concatenated arrays."""
        ends = np.cumsum(self._counts)
        return [slice(int(e - c), int(e)) for c, e in zip(self._counts, ends)]

    def set_scene_colliders(self, scene: int, col: Colliders):
        boxes, nb, spheres, ns, planes, npl = col.c_arrays()
        check(lib().lps_set_scene_colliders(self._ctx, int(scene), C.cast(boxes, C.c_void_p), nb,
                                            C.cast(spheres, C.c_void_p), ns, C.cast(planes, C.c_void_p), npl),
              "lps_set_scene_colliders")

    def scene_stats(self):
        S = self.scene_count
        st = (CStepStats * S)()
        check(lib().lps_get_scene_stats(self._ctx, C.cast(st, C.c_void_p), S), "lps_get_scene_stats")
        return list(st)


def solver_for(params: Params):
    return PCISPHSolver(params) if params.method == METHOD_PCISPH else SPHSolver(params)
