"""bemuse_amd: MI355X (gfx950) implementation of the BEMUse per-frequency
hydrodynamic hot path, driven through the C-ABI of ``libbemuse_amd.so``
(declared in ``include/bemuse_amd.h``).

The Python layer is a thin ctypes binding: every arithmetic step (Rankine and
Havelock influence assembly, complex LU, solves, reductions) runs in the HIP
library on the GPU. There is no CPU fallback: if the library is missing the
import of the binding fails loudly.
"""
from __future__ import annotations

import atexit
import ctypes
import math
import os
from ctypes import POINTER, c_double, c_int, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbemuse_amd.so")

MODE_FP64, MODE_MIXED, MODE_MIXED_SURROGATE = 0, 1, 2
COMPOSE_WAVE, COMPOSE_ZERO, COMPOSE_INF = 0, 1, 2
GRAVITY = 9.81          # BEMUse_Math_Types.h:115
TWO_PI_INV = 0.5 / math.pi

_lib = None
_exiting = False


def _mark_exit():
    # Objects collected after interpreter shutdown starts must not call into
    # the HIP runtime, which may already be torn down.
    global _exiting
    _exiting = True


atexit.register(_mark_exit)


class BemError(RuntimeError):
    """A failed C-ABI call; .code is the BEM_ERR_* status (ERR_* below)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NUMERIC, ERR_STATE = 1, 2, 3, 4, 5
NOALLOC_MAX_N = 512     # BEM_NOALLOC_MAX_N
LU_FP32, LU_SPLIT_F16, LU_SPLIT_F16_TILE = 0, 1, 2


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("refine_iters", c_int),
        ("fell_back_fp64", c_int),
        ("pivot_fallback", c_int),
        ("residual", c_double),
        ("ms_assemble", c_double),
        ("ms_factor", c_double),
        ("ms_solve", c_double),
        ("surrogate_points", ctypes.c_int64),
        ("exact_points", ctypes.c_int64),
        ("surrogate_err", c_double),
        ("lanes", c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SurrogateConfig(ctypes.Structure):
    _fields_ = [
        ("seed", c_uint64),
        ("n_samples", ctypes.c_int64),
        ("batch", c_int),
        ("steps", c_int),
        ("lr", c_double),
        ("lr_final", c_double),
        ("n_valid", ctypes.c_int64),
        ("sampling", c_int),
        ("steps_per_launch", c_int),
    ]


class SurrogateReport(ctypes.Structure):
    _fields_ = [
        ("loss_initial", c_double),
        ("loss_final", c_double),
        ("valid_rms", c_double * 2),
        ("valid_max", c_double * 2),
        ("rel_rms", c_double),
        ("rel_p99", c_double),
        ("rel_max", c_double),
        ("ms_sample", c_double),
        ("ms_train", c_double),
        ("ms_valid", c_double),
    ]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if k.startswith("valid_") else v
        return out


_PD = POINTER(c_double)
_PI = POINTER(c_int)
_PP = POINTER(c_void_p)
_PF = POINTER(ctypes.c_float)

_SIGS = {
    "bem_last_error": (ctypes.c_char_p, []),
    "bem_mesh_from_arrays": (c_int, [c_int, _PD, c_int, _PI, _PP]),
    "bem_mesh_half_cylinder": (c_int, [c_double, c_double, c_int, c_int, c_int, c_int, _PP]),
    "bem_mesh_box": (c_int, [c_double, c_double, c_double, c_double, c_int, c_int, c_int, c_int, _PP]),
    "bem_mesh_semisub3": (c_int, [c_double, c_double, c_double, c_double, c_double, c_double,
                                  c_int, c_int, c_int, c_int, c_int, c_int, _PP]),
    "bem_mesh_cone_spar": (c_int, [c_double, c_double, c_double, c_double, c_double,
                                   c_int, c_int, c_int, c_int, c_int, _PP]),
    "bem_mesh_random_body": (c_int, [c_uint64, c_int, _PD, _PP]),
    "bem_mesh_size": (c_int, [c_void_p, _PI, _PI]),
    "bem_mesh_export": (c_int, [c_void_p, _PD, _PI]),
    "bem_mesh_free": (None, [c_void_p]),
    "bem_geometry": (c_int, [c_void_p, _PD, _PD, _PD, _PD, _PD]),
    "bem_create": (c_int, [c_int, _PP]),
    "bem_destroy": (None, [c_void_p]),
    "bem_body_create": (c_int, [c_void_p, c_void_p, _PP]),
    "bem_body_destroy": (None, [c_void_p]),
    "bem_body_n": (c_int, [c_void_p]),
    "bem_body_n_system": (c_int, [c_void_p]),
    "bem_body_create_lid": (c_int, [c_void_p, c_void_p, c_void_p, _PP]),
    "bem_mesh_lid_disc": (c_int, [c_double, c_int, c_int, c_int, _PP]),
    "bem_mesh_lid_rect": (c_int, [c_double, c_double, c_int, c_int, c_int, _PP]),
    "bem_mesh_append": (c_int, [c_void_p, c_void_p, c_double, c_double, c_double]),
    "bem_mesh_lid_ellipse": (c_int, [c_double, c_double, c_int, c_int, c_int, _PP]),
    "bem_mesh_revolution": (c_int, [c_int, _PD, _PD, c_int, _PP]),
    "bem_mesh_tapered_spar": (c_int, [c_double] * 5 + [c_int] * 6 + [_PP]),
    "bem_mesh_spar_leg": (c_int, [c_double] * 4 + [c_int] * 6 + [_PP]),
    "bem_mesh_barge": (c_int, [c_double] * 3 + [c_int] * 4 + [_PP]),
    "bem_mesh_ellipsoid": (c_int, [c_double] * 4 + [c_int] * 4 + [_PP]),
    "bem_mesh_semi_ellipsoid": (c_int, [c_double] * 3 + [c_int] * 4 + [_PP]),
    "bem_mesh_wigley": (c_int, [c_double] * 3 + [c_int] * 5 + [_PP]),
    "bem_body_assemble_rankine": (c_int, [c_void_p, c_void_p]),
    "bem_solve_frequency": (c_int, [c_void_p, c_void_p, c_int, c_double, c_double, c_int, c_int, c_int, _PD,
                                    _PD, _PD, _PD, _PD, _PD, _PD, _PD, POINTER(SolveInfo)]),
    "bem_get_rankine": (c_int, [c_void_p, c_void_p, _PD, _PD, _PD, _PD]),
    "bem_get_system": (c_int, [c_void_p, c_void_p, c_int, c_double, c_int, _PD, _PD]),
    "bem_eval_havelock": (c_int, [c_void_p, c_int, c_int, _PD, _PD, _PD, _PD]),
    "bem_gf_coefficients": (c_int, [_PD, c_int]),
    "bem_struve_coefficients": (c_int, [_PD, c_int]),
    "bem_bessel01": (c_int, [c_int, _PD, _PD]),
    "bem_mesh_read": (c_int, [ctypes.c_char_p, c_int, ctypes.POINTER(c_void_p)]),
    "bem_solve_batch": (c_int, [c_void_p, c_int, ctypes.POINTER(c_void_p), _PI, _PD, c_int, c_int, _PD, c_int, _PD,
                                _PD, _PD, c_void_p]),
    "bem_dense_solve": (c_int, [c_void_p, c_int, c_int, _PD, c_int, _PD, _PD, POINTER(SolveInfo)]),
    "bem_dense_solve_noalloc": (c_int, [c_void_p, c_int, c_int, _PD, c_int, _PD, _PD, POINTER(SolveInfo)]),
    "bem_profile_lu": (c_int, [c_void_p, c_void_p, c_int, c_double, _PD]),
    "bem_set_graphs": (c_int, [c_void_p, c_int]),
    "bem_set_refine_limit": (c_int, [c_void_p, c_int]),
    "bem_set_lu_arith": (c_int, [c_void_p, c_int]),
    "bem_set_lu_tile": (c_int, [c_void_p, c_int]),
    "bem_set_lu_speculative": (c_int, [c_void_p, c_int]),
    "bem_set_refine_eager": (c_int, [c_void_p, c_int]),
    "bem_set_pivot_threshold": (c_int, [c_void_p, c_double]),
    "bem_set_wave_far32": (c_int, [c_void_p, c_double]),
    "bem_set_wave_wg_per_cu": (c_int, [c_void_p, c_int]),
    "bem_measure_peaks": (c_int, [c_void_p, _PD]),
    "bem_kochin": (c_int, [c_void_p, c_void_p, c_double, c_double, c_int, c_int, _PD, _PD, _PD, _PD, _PD]),
    "bem_fs_elevation": (c_int, [c_void_p, c_void_p, c_double, c_double, c_int, _PD, c_int, _PD, _PD, _PD, _PD,
                                 _PD]),
    "bem_surrogate_default_config": (c_int, [POINTER(SurrogateConfig)]),
    "bem_surrogate_param_count": (c_int, []),
    "bem_surrogate_train": (c_int, [c_void_p, POINTER(SurrogateConfig), POINTER(SurrogateReport)]),
    "bem_surrogate_get": (c_int, [c_void_p, _PF, c_int]),
    "bem_surrogate_set": (c_int, [c_void_p, _PF, c_int]),
    "bem_surrogate_sample": (c_int, [c_void_p, c_uint64, c_int, ctypes.c_int64, c_int, _PF, _PF]),
    "bem_surrogate_forward": (c_int, [c_void_p, _PF, c_int, _PF, _PF]),
    "bem_surrogate_grad": (c_int, [c_void_p, _PF, c_int, _PF, _PF, _PF, _PD]),
    "bem_surrogate_eval": (c_int, [c_void_p, c_int, _PD, _PD, _PD, _PI]),
    "bem_solve_sweep": (c_int, [c_void_p, c_void_p, c_int, c_int, _PD, _PD, c_int, c_int, _PD, c_int, _PD, _PD,
                                _PD, _PD, _PD, POINTER(SolveInfo)]),
}


def exported_symbols():
    """Names of the C-ABI entry points this binding expects the library to export."""
    return sorted(_SIGS)


def lib():
    """Loads libbemuse_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BemError(f"{LIB_PATH} not built: run __graft_entry__.build() (make -C bemuse_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().bem_last_error().decode(errors="replace")
        raise BemError(f"bemuse_amd error {rc}: {msg}", rc)


def _dp(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_PI) if a is not None else None


def _fp(a):
    return a.ctypes.data_as(_PF) if a is not None else None


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def surrogate_param_count():
    return int(lib().bem_surrogate_param_count())


def surrogate_default_config():
    """Default training configuration (dict) of the wave-term surrogate."""
    c = SurrogateConfig()
    _check(lib().bem_surrogate_default_config(ctypes.byref(c)))
    return {k: getattr(c, k) for k, _ in c._fields_}


def bessel01(x):
    """(n, 4) J0, J1, Y0, Y1 from the device code's Bessel evaluator, run on the host (Y = 0 for x <= 3)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((len(x), 4))
    _check(lib().bem_bessel01(len(x), _dp(x), _dp(out)))
    return out


def gf_coefficients():
    """The HavelockGF polynomial tables and EUL the device code is built with (65 values)."""
    out = np.zeros(65)
    _check(lib().bem_gf_coefficients(_dp(out), 65))
    return out


def struve_coefficients():
    """The Struve approximation literals the device code is built with (26 values)."""
    out = np.zeros(26)
    _check(lib().bem_struve_coefficients(_dp(out), 26))
    return out


# ---------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------
class Mesh:
    """Surface panel mesh (nodes n x 3, elements e x 4 with -1 marking triangles)."""

    def __init__(self, handle):
        self._h = c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and not _exiting:
            _lib.bem_mesh_free(self._h)
            self._h = None

    @staticmethod
    def _make(fn, *args):
        h = c_void_p()
        _check(fn(*args, ctypes.byref(h)))
        return Mesh(h.value)

    @classmethod
    def from_arrays(cls, nodes, elems):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
        elems = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 4)
        return cls._make(lib().bem_mesh_from_arrays, len(nodes), _dp(nodes), len(elems), _ip(elems))

    @classmethod
    def read(cls, path, fmt=0):
        """GDF / MAR / PNL / ASCII STL file (fmt 0: by extension)."""
        return cls._make(lib().bem_mesh_read, os.fsencode(path), int(fmt))

    @classmethod
    def half_cylinder(cls, radius, draft, n_radial, n_vertical, n_azimuthal, cosine=True):
        return cls._make(lib().bem_mesh_half_cylinder, radius, draft, n_radial, n_vertical, n_azimuthal, int(cosine))

    @classmethod
    def box(cls, L, W, z0, z1, nx, ny, nz, open_top=True):
        return cls._make(lib().bem_mesh_box, L, W, z0, z1, nx, ny, nz, int(open_top))

    @classmethod
    def lid_disc(cls, radius, n_radial, n_azimuthal, cosine=True):
        """Interior free-surface lid of a volume of revolution (irregular-frequency removal)."""
        return cls._make(lib().bem_mesh_lid_disc, radius, n_radial, n_azimuthal, int(cosine))

    @classmethod
    def lid_rect(cls, L, W, nx, ny, cosine=False):
        """Interior free-surface lid of a barge."""
        return cls._make(lib().bem_mesh_lid_rect, L, W, nx, ny, int(cosine))

    @classmethod
    def lid_ellipse(cls, a, b, n_radial, n_azimuthal, cosine=True):
        """Interior free-surface lid of a semi-ellipsoid."""
        return cls._make(lib().bem_mesh_lid_ellipse, a, b, n_radial, n_azimuthal, int(cosine))

    @classmethod
    def revolution(cls, r, z, n_azimuthal):
        """Volume of revolution from a profile (r_i, z_i) with r_0 = 0 (the reference's Volume_of_Revolution)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        return cls._make(lib().bem_mesh_revolution, len(r), _dp(r), _dp(z), n_azimuthal)

    @classmethod
    def tapered_spar(cls, r_lower, r_upper, draft, h_taper_begin, h_taper, n_radial, n_vert1, n_vert2, n_vert3,
                     n_azimuthal, cosine=True):
        return cls._make(lib().bem_mesh_tapered_spar, r_lower, r_upper, draft, h_taper_begin, h_taper, n_radial,
                         n_vert1, n_vert2, n_vert3, n_azimuthal, int(cosine))

    @classmethod
    def spar_leg(cls, r_lower, r_upper, draft, h_plate, n_base, n_plate_vert, n_plate_horiz, n_vert, n_azimuthal,
                 cosine=True):
        return cls._make(lib().bem_mesh_spar_leg, r_lower, r_upper, draft, h_plate, n_base, n_plate_vert,
                         n_plate_horiz, n_vert, n_azimuthal, int(cosine))

    @classmethod
    def barge(cls, L, W, D, nx, ny, nz, cosine=True):
        """The reference Barge (open-top box, cosine spacing by default)."""
        return cls._make(lib().bem_mesh_barge, L, W, D, nx, ny, nz, int(cosine))

    @classmethod
    def ellipsoid(cls, a, b, c, depth, n_axial, n_azimuthal, cosine=True, triangles=False):
        return cls._make(lib().bem_mesh_ellipsoid, a, b, c, depth, n_axial, n_azimuthal, int(cosine), int(triangles))

    @classmethod
    def semi_ellipsoid(cls, a, b, c, n_axial, n_azimuthal, cosine=True, triangles=False):
        return cls._make(lib().bem_mesh_semi_ellipsoid, a, b, c, n_axial, n_azimuthal, int(cosine), int(triangles))

    @classmethod
    def wigley(cls, L, B, T, nx, nz, cosine=True, triangles=False, outward=False):
        return cls._make(lib().bem_mesh_wigley, L, B, T, nx, nz, int(cosine), int(triangles), int(outward))

    def append(self, other, dx=0.0, dy=0.0, dz=0.0):
        _check(lib().bem_mesh_append(self._h, other._h, dx, dy, dz))
        return self

    @classmethod
    def semisub3(cls, col_radius=6.0, draft=20.0, spacing=50.0, pontoon_width=6.0, pontoon_height=4.0, gap=0.5,
                 n_azimuthal=44, n_vertical=24, n_radial=8, nx=31, ny=5, nz=4):
        return cls._make(lib().bem_mesh_semisub3, col_radius, draft, spacing, pontoon_width, pontoon_height, gap,
                         n_azimuthal, n_vertical, n_radial, nx, ny, nz)

    @classmethod
    def cone_spar(cls, r_top, r_bot, draft, r_plate, t_plate, n_azimuthal, n_disc, n_rim, n_annulus, n_cone):
        return cls._make(lib().bem_mesh_cone_spar, r_top, r_bot, draft, r_plate, t_plate, n_azimuthal, n_disc,
                         n_rim, n_annulus, n_cone)

    @classmethod
    def random_body(cls, seed, index):
        params = np.zeros(6)
        h = c_void_p()
        _check(lib().bem_mesh_random_body(c_uint64(seed), index, _dp(params), ctypes.byref(h)))
        m = Mesh(h.value)
        m.params = dict(type="cylinder" if params[0] == 0 else "barge", a=params[1], b=params[2], draft=params[3],
                        target=int(params[4]), n=int(params[5]))
        return m

    def size(self):
        nn, ne = c_int(), c_int()
        _check(lib().bem_mesh_size(self._h, ctypes.byref(nn), ctypes.byref(ne)))
        return nn.value, ne.value

    @property
    def n_panels(self):
        return self.size()[1]

    def arrays(self):
        nn, ne = self.size()
        nodes = np.zeros((nn, 3))
        elems = np.zeros((ne, 4), dtype=np.int32)
        _check(lib().bem_mesh_export(self._h, _dp(nodes), _ip(elems)))
        return nodes, elems

    def geometry(self):
        """Per-panel host records (see bem_internal.h PanelField / QuadField)."""
        ne = self.n_panels
        body = np.zeros((ne, 32)); refl = np.zeros((ne, 32)); quad = np.zeros((ne, 32))
        recv = np.zeros((ne, 4)); gen = np.zeros((ne, 8))
        _check(lib().bem_geometry(self._h, _dp(body), _dp(refl), _dp(quad), _dp(recv), _dp(gen)))
        return dict(body=body, refl=refl, quad=quad, recv=recv, gen=gen)


# ---------------------------------------------------------------------------
# Context / body / solve
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        h = c_void_p()
        _check(lib().bem_create(device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().bem_destroy(self._h)
            self._h = None

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass

    def eval_havelock(self, X, Y, V, precision=64):
        """HavelockGF on the device; precision 32 evaluates the far-pair FP32 variant."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        out = np.zeros((len(X), 8))
        _check(lib().bem_eval_havelock(self._h, int(precision), len(X), _dp(X), _dp(Y), _dp(V), _dp(out)))
        return out

    def measure_peaks(self):
        """Measured FP32 / FP64 / fp16 MFMA TFLOP/s and streaming-copy GB/s of the device."""
        out = np.zeros(4)
        _check(lib().bem_measure_peaks(self._h, _dp(out)))
        return dict(fp32_mfma_tflops=float(out[0]), fp64_mfma_tflops=float(out[1]), copy_gbs=float(out[2]),
                    fp16_mfma_tflops=float(out[3]))

    def set_refine_limit(self, max_steps):
        """Correction steps allowed to the mixed modes' refinement (0..4, default 4);
        0 sends every mixed frequency to the FP64 re-solve."""
        _check(lib().bem_set_refine_limit(self._h, int(max_steps)))

    def set_refine_eager(self, steps):
        """Correction steps a solve lane enqueues with its frequency (0..4, default 2); the
        rest follow when the refinement needs them. Results do not depend on it."""
        _check(lib().bem_set_refine_eager(self._h, int(steps)))

    def set_wave_far32(self, near_radius):
        """Measurement option: the wave term of pairs farther than near_radius (m) with FP32
        special functions; near_radius < 0 (default) turns it off."""
        _check(lib().bem_set_wave_far32(self._h, float(near_radius)))

    def set_wave_wg_per_cu(self, n):
        """Work-groups per CU of the wave assembly: -1 by order (default: 2 below padded
        order 6144, else 0), 0 one work-group per tile, n > 0 that many per CU taking the
        tiles in turn. Results do not depend on it."""
        _check(lib().bem_set_wave_wg_per_cu(self._h, int(n)))

    def set_lu_arith(self, arith):
        """Arithmetic of the mixed modes' FP32 LU updates: LU_SPLIT_F16 (default), LU_FP32, or
        LU_SPLIT_F16_TILE (the default's arithmetic with the split in every tile)."""
        _check(lib().bem_set_lu_arith(self._h, int(arith)))

    def set_lu_tile(self, tile):
        """Tile of the LU_SPLIT_F16 trailing updates: -1 autotuned (default), 0-2 the
        32x32-wave-block tiles (64x64, 128x64, 64x128) with loads one chunk ahead, 5-7 the
        same two chunks ahead, 3-4 the 128x128 tiles with 64x64 wave blocks (one or two
        LDS chunks), 8 the 128x128 tile with direct memory-to-LDS operand copies. Results do
        not depend on it."""
        _check(lib().bem_set_lu_tile(self._h, int(tile)))

    def set_lu_speculative(self, enable):
        """Solve lanes factor speculatively (default True: no per-panel pivot check, a
        frequency needing interchanges re-run with it) or with the pivot check after every
        panel (False). Results are the same bit for bit wherever no panel is rejected."""
        _check(lib().bem_set_lu_speculative(self._h, int(bool(enable))))

    def set_pivot_threshold(self, tau):
        """Multiplier bound of the pivot check (0..1, default 1): panels with a multiplier
        |l| > tau are factored again with partial pivoting (0: every panel)."""
        _check(lib().bem_set_pivot_threshold(self._h, float(tau)))

    def set_graphs(self, enable):
        """Capture and replay per-frequency HIP graphs on the solve lanes (default off)."""
        _check(lib().bem_set_graphs(self._h, int(bool(enable))))

    # ---- wave-term surrogate ----
    def surrogate_train(self, **overrides):
        """Samples, trains and validates the wave-term surrogate on this
        context's device (keys of surrogate_default_config() override the
        defaults); returns the training report as a dict."""
        c = SurrogateConfig()
        _check(lib().bem_surrogate_default_config(ctypes.byref(c)))
        for k, v in overrides.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
        rep = SurrogateReport()
        _check(lib().bem_surrogate_train(self._h, ctypes.byref(c), ctypes.byref(rep)))
        return rep.as_dict()

    def surrogate_get(self):
        p = np.zeros(surrogate_param_count(), dtype=np.float32)
        _check(lib().bem_surrogate_get(self._h, _fp(p), len(p)))
        return p

    def surrogate_set(self, params):
        p = _f32(params)
        _check(lib().bem_surrogate_set(self._h, _fp(p), len(p)))

    def surrogate_sample(self, seed, first, n, sampling=1):
        X = np.zeros((n, 4), dtype=np.float32)
        T = np.zeros((n, 2), dtype=np.float32)
        _check(lib().bem_surrogate_sample(self._h, seed, sampling, first, n, _fp(X), _fp(T)))
        return X, T

    def surrogate_forward(self, X, params=None):
        X = _f32(X, (-1, 4))
        Y = np.zeros((len(X), 2), dtype=np.float32)
        p = _f32(params) if params is not None else None
        _check(lib().bem_surrogate_forward(self._h, _fp(p), len(X), _fp(X), _fp(Y)))
        return Y

    def surrogate_grad(self, params, X, T):
        p = _f32(params)
        X = _f32(X, (-1, 4))
        T = _f32(T, (-1, 2))
        g = np.zeros(surrogate_param_count(), dtype=np.float32)
        loss = c_double()
        _check(lib().bem_surrogate_grad(self._h, _fp(p), len(X), _fp(X), _fp(T), _fp(g), ctypes.byref(loss)))
        return loss.value, g

    def surrogate_eval(self, h, v):
        h = np.ascontiguousarray(h, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros((len(h), 4))
        used = np.zeros(len(h), dtype=np.int32)
        _check(lib().bem_surrogate_eval(self._h, len(h), _dp(h), _dp(v), _dp(out), _ip(used)))
        return out, used

    def dense_solve(self, A, B, precision=64, noalloc=False):
        """X with A X = B on the device (precision 64: FP64 LU; 32: FP32-MFMA
        LU + FP64 refinement). noalloc: bem_dense_solve_noalloc (no device
        allocation, n <= NOALLOC_MAX_N)."""
        A = np.asfortranarray(A, dtype=np.complex128)
        B = np.asfortranarray(B, dtype=np.complex128)
        if B.ndim == 1:
            B = B[:, None]
        n, nrhs = B.shape
        X = np.zeros((n, nrhs), dtype=np.complex128, order="F")
        info = SolveInfo()
        fn = lib().bem_dense_solve_noalloc if noalloc else lib().bem_dense_solve
        _check(fn(self._h, precision, n, A.ctypes.data_as(_PD), nrhs, B.ctypes.data_as(_PD), X.ctypes.data_as(_PD),
                  ctypes.byref(info)))
        return X, info.as_dict()

    def solve_batch(self, bodies, omegas, betas=(), mode=1, lanes=4):
        """Frequencies of many bodies (omegas: one array per body) on the
        context's shared lanes; returns one dict (A, B, F, info per
        frequency) per body, in order."""
        bodies = list(bodies)
        omegas = [np.ascontiguousarray(w, dtype=np.float64).ravel() for w in omegas]
        assert len(omegas) == len(bodies)
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        nb = len(betas)
        nf = np.array([len(w) for w in omegas], dtype=np.int32)
        allw = np.concatenate(omegas) if len(omegas) else np.zeros(0)
        tot = int(nf.sum())
        A = np.zeros((tot, 36)); B = np.zeros((tot, 36))
        F = np.zeros((tot, 6, max(nb, 1)), dtype=np.complex128)
        infos = (SolveInfo * max(tot, 1))()
        hs = (c_void_p * len(bodies))(*[b._h.value for b in bodies])
        _check(lib().bem_solve_batch(self._h, len(bodies), hs, nf.ctypes.data_as(_PI), _dp(allw), mode, nb,
                                     _dp(betas) if nb else None, lanes, _dp(A), _dp(B),
                                     F.ctypes.data_as(_PD), infos))
        out, o = [], 0
        for k in range(len(bodies)):
            sl = slice(o, o + nf[k])
            out.append(dict(A=A[sl].reshape(-1, 6, 6), B=B[sl].reshape(-1, 6, 6), F=F[sl][:, :, :nb],
                            info=[infos[i].as_dict() for i in range(o, o + nf[k])]))
            o += nf[k]
        return out


class Body:
    def __init__(self, ctx: Context, mesh: Mesh, lid: Mesh = None):
        """lid: interior free-surface panels for irregular-frequency removal."""
        h = c_void_p()
        _check(lib().bem_body_create_lid(ctx._h, mesh._h, lid._h if lid is not None else None, ctypes.byref(h)))
        self._h = h
        self.ctx = ctx
        self.N = lib().bem_body_n(h)                # body panels (rows of the potentials)
        self.N_system = lib().bem_body_n_system(h)  # body + lid panels

    def close(self):
        if self._h:
            lib().bem_body_destroy(self._h)
            self._h = None

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass

    def assemble_rankine(self):
        _check(lib().bem_body_assemble_rankine(self.ctx._h, self._h))

    def get_rankine(self):
        N = self.N_system
        out = [np.zeros((N, N), order="F") for _ in range(4)]
        _check(lib().bem_get_rankine(self.ctx._h, self._h, *[_dp(o) for o in out]))
        return dict(Ssrc=out[0], Srefl=out[1], Dsrc=out[2], Drefl=out[3])

    def get_system(self, kappa, mode=MODE_FP64, compose=COMPOSE_WAVE):
        N = self.N_system
        S = np.zeros((N, N), dtype=np.complex128, order="F")
        D = np.zeros((N, N), dtype=np.complex128, order="F")
        _check(lib().bem_get_system(self.ctx._h, self._h, mode, kappa, compose, S.ctypes.data_as(_PD),
                                    D.ctypes.data_as(_PD)))
        return S, D

    def profile_lu(self, kappa, mode=MODE_FP64):
        """Instrumented factorisation: GEMM launch timings of one LU (see bem_profile_lu)."""
        out = np.zeros(9)
        _check(lib().bem_profile_lu(self.ctx._h, self._h, mode, kappa, _dp(out)))
        return dict(ms_lu=out[0], ms_gemm=out[1], gemm_flops=out[2], gemm_launches=int(out[3]),
                    ms_gemm_per_launch=out[4], flops_per_launch=out[5], gemm_bytes=out[6],
                    bytes_per_launch=out[7], ms_gemm_per_launch_replay=out[8])

    def solve_sweep(self, omegas, betas=(), mode=MODE_FP64, lanes=4, kappas=None, compose=COMPOSE_WAVE):
        """Frequencies pipelined over `lanes` concurrent solve lanes; returns
        stacked A (nf,6,6), B, F/FK/SC (nf,6,nb) and a list of info dicts."""
        omegas = np.ascontiguousarray(omegas, dtype=np.float64)
        nf = len(omegas)
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        nb = len(betas)
        kap = None if kappas is None else np.ascontiguousarray(kappas, dtype=np.float64)
        A = np.zeros((nf, 36)); B = np.zeros((nf, 36))
        F = np.zeros((nf, 6, max(nb, 1)), dtype=np.complex128)
        FK = np.zeros_like(F); SC = np.zeros_like(F)
        infos = (SolveInfo * max(nf, 1))()
        _check(lib().bem_solve_sweep(self.ctx._h, self._h, mode, nf, _dp(omegas), _dp(kap), compose, nb,
                                     _dp(betas) if nb else None, lanes, _dp(A), _dp(B), F.ctypes.data_as(_PD),
                                     FK.ctypes.data_as(_PD), SC.ctypes.data_as(_PD), infos))
        return dict(A=A.reshape(nf, 6, 6), B=B.reshape(nf, 6, 6), F=F[:, :, :nb], FK=FK[:, :, :nb],
                    SC=SC[:, :, :nb], info=[infos[i].as_dict() for i in range(nf)])

    def _potentials(self, phiJ, phiS, nb):
        pj = np.asfortranarray(phiJ, dtype=np.complex128).reshape(self.N, 6, order="F")
        ps = np.asfortranarray(phiS, dtype=np.complex128).reshape(self.N, nb, order="F") if nb else None
        return pj, ps

    def kochin(self, omega, phiJ, phiS=None, betas=(), nkoch=36, kappa=None):
        """Kochin functions of a solved frequency (potentials as solve(..., want_phi=True)
        returns them): (KochRad nkoch x 6, KochDiff nkoch x nb) complex."""
        kappa = omega * omega / 9.81 if kappa is None else kappa
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        nb = len(betas)
        pj, ps = self._potentials(phiJ, phiS, nb)
        rad = np.zeros((nkoch, 6), dtype=np.complex128)
        diff = np.zeros((nkoch, max(nb, 1)), dtype=np.complex128)
        _check(lib().bem_kochin(self.ctx._h, self._h, omega, kappa, nkoch, nb, _dp(betas) if nb else None,
                                pj.ctypes.data_as(_PD), ps.ctypes.data_as(_PD) if nb else None,
                                rad.ctypes.data_as(_PD), diff.ctypes.data_as(_PD)))
        return rad, diff[:, :nb]

    def fs_elevation(self, omega, points, phiJ, phiS=None, betas=(), kappa=None):
        """Radiated (npts x 6) and diffracted (npts x nb) free-surface elevation at exterior points."""
        kappa = omega * omega / 9.81 if kappa is None else kappa
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        nb = len(betas)
        pj, ps = self._potentials(phiJ, phiS, nb)
        rad = np.zeros((len(pts), 6), dtype=np.complex128)
        diff = np.zeros((len(pts), max(nb, 1)), dtype=np.complex128)
        _check(lib().bem_fs_elevation(self.ctx._h, self._h, omega, kappa, len(pts), _dp(pts), nb,
                                      _dp(betas) if nb else None, pj.ctypes.data_as(_PD),
                                      ps.ctypes.data_as(_PD) if nb else None, rad.ctypes.data_as(_PD),
                                      diff.ctypes.data_as(_PD)))
        return rad, diff[:, :nb]

    def solve(self, omega, kappa=None, betas=(), mode=MODE_FP64, compose=COMPOSE_WAVE, rad_only=False,
              want_phi=False):
        """One frequency. omega is the solver's internal Omega (rad/s); kappa
        defaults to Omega^2 / 9.81 (infinite depth)."""
        if kappa is None:
            kappa = omega * omega / GRAVITY
        betas = np.ascontiguousarray(betas, dtype=np.float64)
        nb = len(betas)
        A = np.zeros(36); B = np.zeros(36)
        F = np.zeros((6, nb), dtype=np.complex128)
        FK = np.zeros((6, nb), dtype=np.complex128)
        SC = np.zeros((6, nb), dtype=np.complex128)
        phiJ = np.zeros((self.N, 6), dtype=np.complex128, order="F") if want_phi else None
        phiS = np.zeros((self.N, max(nb, 1)), dtype=np.complex128, order="F") if want_phi else None
        info = SolveInfo()
        _check(lib().bem_solve_frequency(
            self.ctx._h, self._h, mode, omega, kappa, compose, int(rad_only), nb, _dp(betas) if nb else None,
            _dp(A), _dp(B), F.ctypes.data_as(_PD), FK.ctypes.data_as(_PD), SC.ctypes.data_as(_PD),
            phiJ.ctypes.data_as(_PD) if want_phi else None, phiS.ctypes.data_as(_PD) if want_phi else None,
            ctypes.byref(info)))
        res = dict(A=A.reshape(6, 6), B=B.reshape(6, 6), F=F, FK=FK, SC=SC, info=info.as_dict())
        if want_phi:
            res["phiJ"] = phiJ
            res["phiS"] = phiS[:, :nb]
        return res
