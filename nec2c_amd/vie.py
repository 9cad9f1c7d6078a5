"""Python mirror of the reference's VIE FFT-GMRES interface over libvie_hip.so.

The names, argument meaning and error behaviour follow the reference's C API
(vie_fft.h, vie_iterative.h, dielectric_fft.h); every call goes through the
C-ABI of nec2c_amd/lib/libvie_hip.so, whose kernels run on the GPU.  There is
no CPU compute path here: if the library is missing, ``lib()`` raises.

Vectors are numpy complex128 arrays in the reference's layout: conductor
unknowns first, then 3 interleaved components per cell in the caller's cell
order.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvie_hip.so")

FP64, MIXED = 0, 1
KMAX_RESTART = 128  # longest Krylov cycle of the device solver (krylov.h kMaxRestart)
STATUS = ("converged", "iteration limit", "Krylov breakdown",
          "operator/preconditioner failure", "invalid input",
          "workspace allocation/size failure", "nonfinite arithmetic")


class CDouble(ctypes.Structure):
    """C99 ``double _Complex`` passed by value (two SSE doubles on x86-64)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)


VectorMap = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)


class LinearOperator(ctypes.Structure):        # vie_iterative.h:11-18
    _fields_ = [("order", ctypes.c_size_t), ("apply", ctypes.c_void_p), ("context", ctypes.c_void_p),
                ("precondition", ctypes.c_void_p), ("precondition_context", ctypes.c_void_p)]


class GmresOptions(ctypes.Structure):          # vie_iterative.h:20-26
    _fields_ = [("restart", ctypes.c_size_t), ("maximum_iterations", ctypes.c_size_t),
                ("relative_tolerance", ctypes.c_double), ("absolute_tolerance", ctypes.c_double)]


class GmresReport(ctypes.Structure):           # vie_iterative.h:39-47
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_size_t),
                ("operator_applications", ctypes.c_size_t), ("restarts", ctypes.c_size_t),
                ("residual_norm", ctypes.c_double), ("relative_residual", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SolveOptions(ctypes.Structure):
    """vie_hip_solve_options; recycle: recycled subspace size (0 default per
    precision, -1 off)."""
    _fields_ = [("precision", ctypes.c_int), ("inner_reduction", ctypes.c_double), ("recycle", ctypes.c_int)]


class Clearance(ctypes.Structure):             # vie_hip_clearance (include/vie_hip.h)
    _fields_ = [("worst_margin", ctypes.c_double), ("worst_gap", ctypes.c_double),
                ("worst_required", ctypes.c_double), ("minimum_gap", ctypes.c_double),
                ("required_gap", ctypes.c_double), ("worst_box", ctypes.c_longlong),
                ("worst_conductor", ctypes.c_longlong), ("worst_is_patch", ctypes.c_int),
                ("has_conductors", ctypes.c_int), ("insufficient", ctypes.c_int)]


class Sample(ctypes.Structure):                # dielectric.h:18-24
    _fields_ = [("material_id", ctypes.c_int), ("conductivity_kind", ctypes.c_int),
                ("frequency_mhz", ctypes.c_double), ("epsilon_r", ctypes.c_double), ("loss", ctypes.c_double),
                ("source_line", ctypes.c_size_t)]


class Phase(ctypes.Structure):                 # vie_hip.h vie_hip_phase
    _fields_ = [("name", ctypes.c_char_p), ("seconds", ctypes.c_double), ("calls", ctypes.c_ulonglong)]


class Material(ctypes.Structure):              # dielectric.h:26-35
    _fields_ = [("id", ctypes.c_int), ("epsilon_r", ctypes.c_double), ("source_line", ctypes.c_size_t),
                ("epsilon_loss", ctypes.c_double), ("conductivity", ctypes.c_double),
                ("tabulated", ctypes.c_int), ("samples", ctypes.c_void_p), ("sample_count", ctypes.c_size_t)]


class Model(ctypes.Structure):                 # dielectric.h:88-107
    _fields_ = [("loaded", ctypes.c_int), ("format_version", ctypes.c_int),
                ("path_is_absolute", ctypes.c_int), ("mesh_record_present", ctypes.c_int),
                ("resource_path", ctypes.c_char_p),
                ("materials", ctypes.POINTER(Material)), ("material_count", ctypes.c_size_t),
                ("samples", ctypes.c_void_p), ("sample_count", ctypes.c_size_t),
                ("regions", ctypes.c_void_p), ("region_count", ctypes.c_size_t),
                ("boxes", ctypes.c_void_p), ("box_count", ctypes.c_size_t),
                ("max_kh", ctypes.c_double), ("max_cells", ctypes.c_size_t), ("max_aspect", ctypes.c_double)]


class Diagnostic(ctypes.Structure):            # dielectric.h:109-116
    _fields_ = [("resource", ctypes.c_char * (255 * 4 + 1)), ("record", ctypes.c_char * 32),
                ("value", ctypes.c_char * 128), ("message", ctypes.c_char * 256), ("line", ctypes.c_size_t)]


class MeshCell(ctypes.Structure):              # dielectric.h:118-127
    _fields_ = [("box_id", ctypes.c_size_t), ("region_id", ctypes.c_int), ("material_id", ctypes.c_int),
                ("center", ctypes.c_double * 3), ("size", ctypes.c_double * 3),
                ("q", ctypes.c_double), ("aspect", ctypes.c_double)]


class MeshResult(ctypes.Structure):            # dielectric.h:147-178
    _fields_ = [("built", ctypes.c_int), ("accuracy_warning", ctypes.c_int), ("frequency_mhz", ctypes.c_double),
                ("regions", ctypes.c_void_p), ("region_count", ctypes.c_size_t),
                ("boxes", ctypes.c_void_p), ("box_count", ctypes.c_size_t),
                ("cells", ctypes.POINTER(MeshCell)), ("cell_count", ctypes.c_size_t)] + \
               [(f, ctypes.c_size_t) for f in ("vacuum_box_count", "feature_warning_count")] + \
               [(f, ctypes.c_double) for f in ("declared_volume", "active_volume", "vacuum_volume",
                                               "numerical_volume", "volume_relative_error", "q_min",
                                               "q_median", "q_p95", "q_max", "aspect_max")] + \
               [("has_conductors", ctypes.c_int), ("minimum_conductor_gap", ctypes.c_double),
                ("required_conductor_gap", ctypes.c_double)] + \
               [(f, ctypes.c_size_t) for f in ("conductor_unknowns", "volume_unknowns", "total_unknowns",
                                               "matrix_bytes", "workspace_bytes")]


# every symbol the public headers declare (include/*.h)
EXPORTS = (
    "vie_gmres_default_options", "vie_iterative_status_name", "vie_gmres",
    "vie_fft_backend_set", "vie_fft_fftw_available", "vie_fft_backend_requested_name",
    "vie_fft_backend_name", "vie_fft_backend_last_used_name", "vie_fft_backend_clear_last_used",
    "vie_fft_create", "vie_fft_set_permittivity", "vie_fft_destroy", "vie_fft_bytes",
    "vie_fft_threads", "vie_fft_parallel_available", "vie_fft_operator",
    "dielectric_gmres_settings", "dielectric_fft_prepare", "dielectric_fft_prepare_pec",
    "dielectric_fft_free", "dielectric_fft_entry", "dielectric_fft_get_operator",
    "dielectric_fft_storage", "dielectric_fft_threads", "dielectric_fft_backend",
    "dielectric_fft_solve", "vie_fft_solve", "vie_hip_set_solve_options", "vie_hip_get_solve_options",
    "dielectric_fft_budget", "dielectric_fft_set_budget", "dielectric_fft_blocks_changed",
    "dielectric_fft_plan", "vie_hip_last_error", "vie_hip_device_count",
    "dielectric_fft_upload_rhs", "dielectric_fft_solve_resident", "dielectric_fft_download_solution",
    "dielectric_fft_time_apply", "dielectric_fft_apply_bytes", "vie_fft_set_permittivities",
    "vie_hip_synchronize", "vie_hip_copy_probe", "dielectric_fft_apply_ex", "dielectric_fft_apply_noalloc",
    "dielectric_fft_apply_noalloc_bytes", "vie_hip_phases", "vie_hip_phases_reset", "vie_fft_set_variants", "vie_hip_ground_table", "dielectric_fft_prepare_ground", "vie_hip_voxels_load", "vie_hip_voxels_free",
    "vie_hip_voxels_info", "vie_hip_voxels_cells", "dielectric_fft_prepare_voxels", "dielectric_fft_reserve",
    "vie_hip_batch_create", "vie_hip_batch_destroy", "vie_hip_batch_job_bytes", "vie_hip_batch_reserve",
    "vie_hip_batch_submit", "vie_hip_batch_progress", "vie_hip_batch_running", "vie_hip_batch_wait",
    "vie_hip_batch_wait_all", "vie_hip_batch_result", "vie_hip_batch_release", "vie_hip_batch_group_stats",
    "vie_hip_batch_share_subspaces", "vie_hip_batch_shared_stats",
    "vie_fft_tune",
    "vie_hip_memory_info", "dielectric_fft_time_arnoldi", "dielectric_fft_solve_stats",
    "dielectric_fft_last_solve_escalated", "dielectric_fft_last_solve_recycled", "vie_hip_harmonic_ritz",
    "vie_hip_radiation_amplitudes", "vie_hip_scattered_power", "vie_hip_volume_conductor_rows",
    "vie_hip_time_volume_conductor_rows", "vie_hip_conductor_volume_columns",
    "vie_hip_time_conductor_volume_columns", "dielectric_fft_assemble_conductor_rows",
    "dielectric_fft_assemble_conductor_columns", "vie_hip_total_far_field", "vie_hip_total_radiated_power",
    "vie_hip_release_workspaces", "vie_hip_box_overlap", "vie_hip_time_box_overlap", "vie_hip_conductor_clearance",
    "vie_hip_time_conductor_clearance",
)

_LIB = None


def lib() -> ctypes.CDLL:
    """Load libvie_hip.so; raise loudly when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m nec2c_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, d, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
    szp = ctypes.POINTER(ctypes.c_size_t)
    sig = {
        "vie_gmres_default_options": (None, [ctypes.POINTER(GmresOptions)]),
        "vie_iterative_status_name": (ctypes.c_char_p, [i]),
        "vie_gmres": (i, [ctypes.POINTER(LinearOperator), vp, vp, ctypes.POINTER(GmresOptions),
                          ctypes.POINTER(GmresReport)]),
        "vie_fft_backend_set": (i, [ctypes.c_char_p]),
        "vie_fft_fftw_available": (i, []),
        "vie_fft_backend_requested_name": (ctypes.c_char_p, []),
        "vie_fft_backend_name": (ctypes.c_char_p, [vp]),
        "vie_fft_backend_last_used_name": (ctypes.c_char_p, []),
        "vie_fft_backend_clear_last_used": (None, []),
        "vie_fft_create": (i, [ctypes.POINTER(vp), szp, sz, szp, d, d, CDouble, sz]),
        "vie_fft_set_permittivity": (i, [vp, sz, CDouble]),
        "vie_fft_destroy": (None, [vp]),
        "vie_fft_bytes": (sz, [vp]),
        "vie_fft_threads": (sz, [vp]),
        "vie_fft_parallel_available": (i, []),
        "vie_fft_operator": (None, [vp, ctypes.POINTER(LinearOperator)]),
        "dielectric_fft_prepare": (i, [ctypes.POINTER(vp), ctypes.POINTER(Model), ctypes.POINTER(MeshResult),
                                       sz, d, sz, ctypes.POINTER(Diagnostic)]),
        "dielectric_fft_prepare_pec": (i, [ctypes.POINTER(vp), ctypes.POINTER(Model), ctypes.POINTER(MeshResult),
                                           sz, d, sz, ctypes.POINTER(Diagnostic)]),
        "dielectric_fft_free": (None, [vp]),
        "dielectric_fft_entry": (vp, [vp, sz, sz]),
        "dielectric_fft_get_operator": (None, [vp, ctypes.POINTER(LinearOperator)]),
        "dielectric_fft_storage": (sz, [vp]),
        "dielectric_fft_threads": (sz, [vp]),
        "dielectric_fft_backend": (ctypes.c_char_p, [vp]),
        "dielectric_fft_solve": (i, [vp, vp, vp, ctypes.POINTER(GmresOptions), ctypes.POINTER(GmresReport)]),
        "vie_fft_solve": (i, [vp, vp, vp, ctypes.POINTER(GmresOptions), ctypes.POINTER(GmresReport)]),
        "vie_hip_set_solve_options": (None, [ctypes.POINTER(SolveOptions)]),
        "vie_hip_get_solve_options": (None, [ctypes.POINTER(SolveOptions)]),
        "dielectric_fft_budget": (sz, []),
        "dielectric_fft_set_budget": (None, [sz]),
        "dielectric_fft_blocks_changed": (None, [vp]),
        "dielectric_fft_plan": (vp, [vp]),
        "vie_hip_last_error": (ctypes.c_char_p, []),
        "vie_hip_device_count": (i, []),
        "dielectric_fft_upload_rhs": (i, [vp, vp]),
        "dielectric_fft_solve_resident": (i, [vp, ctypes.POINTER(GmresOptions), ctypes.POINTER(SolveOptions),
                                              ctypes.POINTER(GmresReport)]),
        "dielectric_fft_download_solution": (i, [vp, vp]),
        "dielectric_fft_time_apply": (i, [vp, i, i, ctypes.POINTER(d), ctypes.POINTER(d)]),
        "dielectric_fft_apply_bytes": (i, [vp, i, ctypes.POINTER(d), ctypes.POINTER(d)]),
        "vie_fft_set_permittivities": (i, [vp, vp, sz]),
        "vie_hip_synchronize": (i, []),
        "dielectric_fft_apply_ex": (i, [vp, i, i, vp, vp]),
        "dielectric_fft_apply_noalloc": (i, [vp, i, i, vp, vp]),
        "dielectric_fft_apply_noalloc_bytes": (sz, [vp, i]),
        "vie_hip_phases": (sz, [vp, sz]),
        "vie_hip_phases_reset": (None, []),
        "vie_fft_set_variants": (i, [vp, vp]),
        "dielectric_fft_prepare_ground": (i, [vp, vp, vp, sz, ctypes.c_double, sz, ctypes.c_double,
                                              ctypes.c_double, vp]),
        "vie_hip_ground_table": (i, [i, i, i, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, i, vp]),
        "vie_hip_voxels_load": (i, [vp, ctypes.c_char_p, vp]),
        "vie_hip_voxels_free": (None, [vp]),
        "vie_hip_voxels_info": (i, [vp, vp, vp, vp, vp]),
        "vie_hip_voxels_cells": (i, [vp, vp, vp, vp]),
        "dielectric_fft_prepare_voxels": (i, [vp, vp, sz, ctypes.c_double, sz, i, vp]),
        "dielectric_fft_reserve": (i, [vp, ctypes.POINTER(GmresOptions)]),
        "vie_hip_copy_probe": (d, [sz, i]),
        "vie_hip_batch_create": (vp, []),
        "vie_hip_batch_destroy": (None, [vp]),
        "vie_hip_batch_job_bytes": (sz, [vp, ctypes.POINTER(GmresOptions), ctypes.POINTER(SolveOptions)]),
        "vie_hip_batch_reserve": (i, [vp, sz, vp, ctypes.POINTER(GmresOptions)]),
        "vie_hip_batch_submit": (ctypes.c_longlong, [vp, vp, vp, ctypes.POINTER(GmresOptions),
                                                      ctypes.POINTER(SolveOptions)]),
        "vie_hip_batch_progress": (sz, [vp, ctypes.POINTER(ctypes.c_longlong), sz, szp]),
        "vie_hip_batch_running": (sz, [vp]),
        "vie_hip_batch_wait": (i, [vp, ctypes.c_longlong]),
        "vie_hip_batch_wait_all": (i, [vp]),
        "vie_hip_batch_result": (i, [vp, ctypes.c_longlong, vp, ctypes.POINTER(GmresReport),
                                     ctypes.POINTER(d)]),
        "vie_hip_batch_release": (i, [vp, ctypes.c_longlong]),
        "vie_hip_batch_group_stats": (i, [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]),
        "vie_hip_batch_share_subspaces": (i, [vp, i]),
        "vie_hip_batch_shared_stats": (i, [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)]),
        "vie_fft_tune": (i, [vp, ctypes.POINTER(ctypes.c_int)]),
        "vie_hip_memory_info": (i, [szp, szp]),
        "dielectric_fft_time_arnoldi": (i, [vp, i, i, i] + [ctypes.POINTER(d)] * 4),
        "dielectric_fft_solve_stats": (i, [vp, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_ulonglong), i]),
        "dielectric_fft_last_solve_escalated": (i, [vp]),
        "dielectric_fft_last_solve_recycled": (i, [vp]),
        "vie_hip_harmonic_ritz": (i, [vp, i, i, vp]),
        "vie_hip_radiation_amplitudes": (i, [sz, vp, vp, vp, d, i, vp, vp, sz, vp, vp]),
        "vie_hip_scattered_power": (i, [sz, vp, vp, vp, d, i, ctypes.POINTER(d)]),
        "vie_hip_volume_conductor_rows": (i, [d, sz, vp, vp, i, vp, vp, i, vp, vp, vp, vp, i, vp, sz]),
        "vie_hip_time_volume_conductor_rows": (i, [d, sz, vp, vp, i, vp, vp, i, vp, vp, vp, vp, i, i,
                                                   ctypes.POINTER(ctypes.c_float)]),
        "vie_hip_conductor_volume_columns": (i, [d, d, sz, vp, vp, vp, i] + [vp] * 8 + [i, vp, vp, vp, vp, sz, vp, d,
                                                                                     i, vp]),
        "dielectric_fft_assemble_conductor_rows": (i, [vp, d, vp, vp, i, vp, vp, i, vp, vp, vp, vp, i]),
        "vie_hip_release_workspaces": (i, []),
        "vie_hip_box_overlap": (i, [vp, vp, sz, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
        "vie_hip_time_box_overlap": (i, [vp, vp, sz, i, ctypes.POINTER(ctypes.c_longlong),
                                         ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(d)]),
        "vie_hip_conductor_clearance": (i, [vp, vp, vp, sz, i] + [vp] * 8 + [i, vp, vp, ctypes.POINTER(Clearance)]),
        "vie_hip_time_conductor_clearance": (i, [vp, vp, vp, sz, i] + [vp] * 8 +
                                             [i, vp, vp, i, ctypes.POINTER(Clearance), ctypes.POINTER(d)]),
        "vie_hip_total_far_field": (i, [d, d, i] + [vp] * 8 + [i, vp, vp, vp, sz, vp, vp, vp, i, sz, vp, vp, vp, vp]),
        "vie_hip_total_radiated_power": (i, [d, d, i] + [vp] * 8 + [i, vp, vp, vp, sz, vp, vp, vp, i, sz,
                                                                     ctypes.POINTER(d)]),
        "dielectric_fft_assemble_conductor_columns": (i, [vp, d, d, vp, vp, vp, i] + [vp] * 8 +
                                                      [i, vp, vp, vp, vp, vp, d, i]),
        "vie_hip_time_conductor_volume_columns": (i, [d, d, sz, vp, vp, vp, i] + [vp] * 8 +
                                                  [i, vp, vp, vp, vp, sz, vp, d, i, i,
                                                   ctypes.POINTER(ctypes.c_float)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def harmonic_ritz(hbar, k):
    """vie_hip_harmonic_ritz: the k harmonic Ritz vectors of smallest modulus
    of an (m+1) x m Arnoldi Hessenberg, as an m x k array (host computation)."""
    hbar = np.asfortranarray(hbar, dtype=np.complex128)
    m = hbar.shape[1]
    if hbar.shape != (m + 1, m):
        raise ValueError("hbar must be (m+1) x m")
    p = np.zeros((m, k), dtype=np.complex128, order="F")
    if not lib().vie_hip_harmonic_ritz(hbar.ctypes.data, m, k, p.ctypes.data):
        raise RuntimeError("harmonic Ritz vectors failed")
    return p


def _cells_arrays(centers, sizes, current):
    c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
    s = np.ascontiguousarray(sizes, dtype=np.float64).reshape(-1, 3)
    j = np.ascontiguousarray(current, dtype=np.complex128).reshape(-1)
    if len(s) != len(c) or len(j) != 3 * len(c):
        raise ValueError("centers, sizes and current must describe the same cells")
    return c, s, j


def radiation_amplitudes(centers, sizes, current, k0, directions, ground=0, rs=None, rp=None):
    """vie_hip_radiation_amplitudes: [D, 3] scattered far-field amplitudes of a
    volume current (free space, PEC plane, or per-direction rs / rp)."""
    c, s, j = _cells_arrays(centers, sizes, current)
    d = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(d), 3), dtype=np.complex128)
    rsa = rpa = None
    if ground == 2:
        rsa = np.ascontiguousarray(rs, dtype=np.complex128).reshape(-1)
        rpa = np.ascontiguousarray(rp, dtype=np.complex128).reshape(-1)
        if len(rsa) != len(d) or len(rpa) != len(d):
            raise ValueError("one rs and one rp per direction")
    ok = lib().vie_hip_radiation_amplitudes(len(c), c.ctypes.data, s.ctypes.data, j.ctypes.data, float(k0),
                                            int(ground), None if rsa is None else rsa.ctypes.data,
                                            None if rpa is None else rpa.ctypes.data, len(d), d.ctypes.data,
                                            out.ctypes.data)
    if not ok:
        raise RuntimeError(f"radiation amplitudes failed: {last_error()}")
    return out


def scattered_power(centers, sizes, current, k0, ground=0):
    """vie_hip_scattered_power: the reference's 974-direction scattered power (W)."""
    c, s, j = _cells_arrays(centers, sizes, current)
    p = ctypes.c_double()
    if not lib().vie_hip_scattered_power(len(c), c.ctypes.data, s.ctypes.data, j.ctypes.data, float(k0), int(ground),
                                         ctypes.byref(p)):
        raise RuntimeError(f"scattered power failed: {last_error()}")
    return p.value


def _conductor_arrays(centers, sizes, wire_points, wire_tangents, patch_points, patch_t1, patch_t2, patch_scale):
    def arr(x, width):
        a = np.ascontiguousarray(np.zeros((0, width)) if x is None else x, dtype=np.float64)
        return a.reshape(-1, width) if width > 1 else a.reshape(-1)
    c, s = arr(centers, 3), arr(sizes, 3)
    wp, wt = arr(wire_points, 3), arr(wire_tangents, 3)
    pp, t1, t2, ps = arr(patch_points, 3), arr(patch_t1, 3), arr(patch_t2, 3), arr(patch_scale, 1)
    if len(s) != len(c) or len(wt) != len(wp) or not (len(t1) == len(t2) == len(ps) == len(pp)):
        raise ValueError("inconsistent cell, wire or patch arrays")
    return c, s, wp, wt, pp, t1, t2, ps


def volume_conductor_rows(k0, centers, sizes, wire_points=None, wire_tangents=None, patch_points=None,
                          patch_t1=None, patch_t2=None, patch_scale=None, ground=0):
    """vie_hip_volume_conductor_rows: the [wires + 2 patches, 3 cells] rows of
    the conductor equations over the volume unknowns (dielectric_wire.c:1164-1262)."""
    c, s, wp, wt, pp, t1, t2, ps = _conductor_arrays(centers, sizes, wire_points, wire_tangents, patch_points,
                                                     patch_t1, patch_t2, patch_scale)
    out = np.zeros((len(wp) + 2 * len(pp), 3 * len(c)), dtype=np.complex128)
    ok = lib().vie_hip_volume_conductor_rows(float(k0), len(c), c.ctypes.data, s.ctypes.data, len(wp), wp.ctypes.data,
                                             wt.ctypes.data, len(pp), pp.ctypes.data, t1.ctypes.data,
                                             t2.ctypes.data, ps.ctypes.data, int(ground), out.ctypes.data,
                                             out.shape[1])
    if not ok:
        raise RuntimeError(f"volume-to-conductor rows failed: {last_error()}")
    return out


def time_volume_conductor_rows(k0, centers, sizes, wire_points=None, wire_tangents=None, patch_points=None,
                               patch_t1=None, patch_t2=None, patch_scale=None, ground=0, repeats=5):
    """Mean device time (ms) of one pass of the row kernel over resident inputs."""
    c, s, wp, wt, pp, t1, t2, ps = _conductor_arrays(centers, sizes, wire_points, wire_tangents, patch_points,
                                                     patch_t1, patch_t2, patch_scale)
    ms = ctypes.c_float()
    ok = lib().vie_hip_time_volume_conductor_rows(float(k0), len(c), c.ctypes.data, s.ctypes.data, len(wp),
                                                  wp.ctypes.data, wt.ctypes.data, len(pp), pp.ctypes.data,
                                                  t1.ctypes.data, t2.ctypes.data, ps.ctypes.data, int(ground),
                                                  int(repeats), ctypes.byref(ms))
    if not ok:
        raise RuntimeError(f"volume-to-conductor row timing failed: {last_error()}")
    return ms.value


def _columns_arrays(centers, sizes, chi, segments, patches, sources):
    c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
    s = np.ascontiguousarray(sizes, dtype=np.float64).reshape(-1, 3)
    ch = np.ascontiguousarray(np.broadcast_to(np.asarray(chi, dtype=np.complex128), (len(c),)))
    seg = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in segments]
    n = len(seg[0]) if seg else 0
    if len(seg) != 8 or any(len(a) != n for a in seg) or len(s) != len(c):
        raise ValueError("segments: x, y, z, si, bi, cab, sab, salp of equal length; one size per cell")
    if patches is None:
        pat = [np.zeros((0, 3)), np.zeros(0), np.zeros((0, 3)), np.zeros((0, 3))]
    else:
        pat = [np.ascontiguousarray(a, dtype=np.float64) for a in patches]
    m = len(pat[1].reshape(-1))
    stride = 3 * n + 3 * m
    src = np.ascontiguousarray(sources, dtype=np.complex128).reshape(-1, stride) if stride else \
        np.zeros((0, 0), dtype=np.complex128)
    return c, s, ch, seg, n, pat, m, src


def conductor_volume_columns(k0, wlam, centers, sizes, chi, segments, sources, patches=None, rkh=1.0, ground=0):
    """vie_hip_conductor_volume_columns: [columns, 3 cells] volume rows of the
    conductor columns (dielectric_wire.c:316-508).  segments = (x, y, z, si,
    bi, cab, sab, salp) in wavelengths; patches = (points [m, 3], areas [m],
    t1 [m, 3], t2 [m, 3]) in wavelengths or None; sources [columns, 3 n + 3 m]."""
    c, s, ch, seg, n, pat, m, src = _columns_arrays(centers, sizes, chi, segments, patches, sources)
    out = np.zeros((len(src), 3 * len(c)), dtype=np.complex128)
    ok = lib().vie_hip_conductor_volume_columns(float(k0), float(wlam), len(c), c.ctypes.data, s.ctypes.data,
                                                ch.ctypes.data, n, *[a.ctypes.data for a in seg], m,
                                                *[a.ctypes.data for a in pat], len(src), src.ctypes.data,
                                                float(rkh), int(ground), out.ctypes.data)
    if not ok:
        raise RuntimeError(f"conductor-to-volume coupling failed: {last_error()}")
    return out


def time_conductor_volume_columns(k0, wlam, centers, sizes, chi, segments, sources, patches=None, rkh=1.0, ground=0,
                                  repeats=3):
    """Mean device time (ms) of one pass of the coupling kernels over resident inputs."""
    c, s, ch, seg, n, pat, m, src = _columns_arrays(centers, sizes, chi, segments, patches, sources)
    ms = ctypes.c_float()
    ok = lib().vie_hip_time_conductor_volume_columns(float(k0), float(wlam), len(c), c.ctypes.data, s.ctypes.data,
                                                     ch.ctypes.data, n, *[a.ctypes.data for a in seg], m,
                                                     *[a.ctypes.data for a in pat], len(src), src.ctypes.data,
                                                     float(rkh), int(ground), int(repeats), ctypes.byref(ms))
    if not ok:
        raise RuntimeError(f"conductor-to-volume timing failed: {last_error()}")
    return ms.value


def _total_args(segments, segment_current, patches, patch_current, centers, sizes, volume_current):
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    c128 = lambda a: np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
    seg = [f64(a) for a in segments] if segments is not None else [np.zeros(0)] * 7
    n = len(seg[0])
    if len(seg) != 7 or any(len(a) != n for a in seg):
        raise ValueError("segments: x, y, z, si, cab, sab, salp of equal length")
    scur = c128(segment_current) if n else np.zeros(0, complex)
    pts, area = (f64(patches[0]), f64(patches[1])) if patches is not None else (np.zeros(0), np.zeros(0))
    m = len(area)
    pcur = c128(patch_current) if m else np.zeros(0, complex)
    if len(scur) != 3 * n or len(pts) != 3 * m or len(pcur) != 3 * m:
        raise ValueError("three current coefficients per segment and per patch, three coordinates per patch")
    if centers is None:
        c = s = np.zeros(0)
        j = np.zeros(0, complex)
    else:
        c, s, j = _cells_arrays(centers, sizes, volume_current)
    return seg, n, scur, pts, area, m, pcur, c, s, j


def total_far_field(k0, wlam, theta, phi, segments=None, segment_current=None, patches=None, patch_current=None,
                    centers=None, sizes=None, volume_current=None, ground=0):
    """vie_hip_total_far_field: (r E_theta, r E_phi) of wires, patches and volume."""
    seg, n, scur, pts, area, m, pcur, c, s, j = _total_args(segments, segment_current, patches, patch_current,
                                                            centers, sizes, volume_current)
    th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
    ph = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
    et = np.zeros(len(th), dtype=np.complex128)
    ep = np.zeros(len(th), dtype=np.complex128)
    ok = lib().vie_hip_total_far_field(float(k0), float(wlam), n, *[a.ctypes.data for a in seg], scur.ctypes.data, m,
                                       pts.ctypes.data, area.ctypes.data, pcur.ctypes.data, len(c) // 3 if c.ndim == 1
                                       else len(c), c.ctypes.data, s.ctypes.data, j.ctypes.data, int(ground), len(th),
                                       th.ctypes.data, ph.ctypes.data, et.ctypes.data, ep.ctypes.data)
    if not ok:
        raise RuntimeError(f"total far field failed: {last_error()}")
    return et, ep


def total_radiated_power(k0, wlam, count=3896, segments=None, segment_current=None, patches=None,
                         patch_current=None, centers=None, sizes=None, volume_current=None, ground=0):
    """vie_hip_total_radiated_power: dw_radiated_power of the coupled solve (W)."""
    seg, n, scur, pts, area, m, pcur, c, s, j = _total_args(segments, segment_current, patches, patch_current,
                                                            centers, sizes, volume_current)
    p = ctypes.c_double()
    ok = lib().vie_hip_total_radiated_power(float(k0), float(wlam), n, *[a.ctypes.data for a in seg],
                                            scur.ctypes.data, m, pts.ctypes.data, area.ctypes.data, pcur.ctypes.data,
                                            len(c) // 3 if c.ndim == 1 else len(c), c.ctypes.data, s.ctypes.data,
                                            j.ctypes.data, int(ground), int(count), ctypes.byref(p))
    if not ok:
        raise RuntimeError(f"total radiated power failed: {last_error()}")
    return p.value


def _corners(minimum, maximum):
    mn = np.ascontiguousarray(minimum, dtype=np.float64).reshape(-1, 3)
    mx = np.ascontiguousarray(maximum, dtype=np.float64).reshape(-1, 3)
    if mn.shape != mx.shape:
        raise ValueError("minimum and maximum corners differ in count")
    return mn, mx


def box_overlap(minimum, maximum, repeats=0):
    """vie_hip_box_overlap: 0-based (left, right) of the first pair of boxes
    whose interiors overlap in validate_box_relationships' loop order
    (dielectric_mesh.c:81-123), or None.  repeats > 0 times the scan instead
    and returns (pair, ms)."""
    mn, mx = _corners(minimum, maximum)
    a, b = ctypes.c_longlong(), ctypes.c_longlong()
    if repeats:
        ms = ctypes.c_double()
        ok = lib().vie_hip_time_box_overlap(mn.ctypes.data, mx.ctypes.data, len(mn), int(repeats), ctypes.byref(a),
                                            ctypes.byref(b), ctypes.byref(ms))
    else:
        ok = lib().vie_hip_box_overlap(mn.ctypes.data, mx.ctypes.data, len(mn), ctypes.byref(a), ctypes.byref(b))
    if not ok:
        raise RuntimeError(f"box overlap failed: {last_error()}")
    pair = None if a.value < 0 else (a.value, b.value)
    return (pair, ms.value) if repeats else pair


def conductor_clearance(minimum, maximum, cell_longest, segments=None, patches=None, repeats=0):
    """vie_hip_conductor_clearance: evaluate_clearance (dielectric_mesh.c:
    215-316) of boxes against NEC wire segments (x, y, z, si, cab, sab, salp,
    bi) and patches (points M x 3, areas).  Returns a dict of the Clearance
    fields; repeats > 0 also returns the average ms."""
    mn, mx = _corners(minimum, maximum)
    cl = np.ascontiguousarray(cell_longest, dtype=np.float64).reshape(-1)
    if len(cl) != len(mn):
        raise ValueError("one cell_longest per box")
    seg = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (segments or [[]] * 8)]
    if len(seg) != 8 or len({len(a) for a in seg}) != 1:
        raise ValueError("segments are eight arrays of one length")
    pts = np.ascontiguousarray(patches[0] if patches else np.zeros((0, 3)), dtype=np.float64).reshape(-1, 3)
    area = np.ascontiguousarray(patches[1] if patches else np.zeros(0), dtype=np.float64).reshape(-1)
    if len(pts) != len(area):
        raise ValueError("one area per patch point")
    out = Clearance()
    args = [mn.ctypes.data, mx.ctypes.data, cl.ctypes.data, len(mn), len(seg[0])] + [a.ctypes.data for a in seg] + \
        [len(area), pts.ctypes.data, area.ctypes.data]
    if repeats:
        ms = ctypes.c_double()
        ok = lib().vie_hip_time_conductor_clearance(*args, int(repeats), ctypes.byref(out), ctypes.byref(ms))
    else:
        ok = lib().vie_hip_conductor_clearance(*args, ctypes.byref(out))
    if not ok:
        raise RuntimeError(f"conductor clearance failed: {last_error()}")
    r = {k: getattr(out, k) for k, _ in Clearance._fields_}
    return (r, ms.value) if repeats else r


def phases(reset=False):
    """The library's timing phases: [{"name", "seconds", "calls"}] (vie_hip_phases)."""
    L = lib()
    n = L.vie_hip_phases(None, 0)
    arr = (Phase * max(n, 1))()
    n = L.vie_hip_phases(arr, n)
    out = [{"name": arr[i].name.decode(), "seconds": arr[i].seconds, "calls": int(arr[i].calls)} for i in range(n)]
    if reset:
        L.vie_hip_phases_reset()
    return out


def ground_table(shape, h, z0, k0, eps, order=48):
    """vie_hip_ground_table: reflected tensor [6, nx, ny, 2 nz - 1] (xx, yy, zz, xy, xz, yz)."""
    nx, ny, nz = shape
    out = np.zeros(6 * nx * ny * (2 * nz - 1), dtype=np.complex128)
    eps = complex(eps)
    if not lib().vie_hip_ground_table(nx, ny, nz, float(h), float(z0), float(k0), eps.real, eps.imag, int(order),
                                      out.ctypes.data):
        raise RuntimeError(f"ground table failed: {last_error()}")
    return out.reshape(6, nx, ny, 2 * nz - 1)


def last_error() -> str:
    return lib().vie_hip_last_error().decode()


def device_count() -> int:
    return lib().vie_hip_device_count()


def memory_info():
    """(free, total) bytes of the current device."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    if not lib().vie_hip_memory_info(ctypes.byref(f), ctypes.byref(t)):
        raise RuntimeError("no HIP device")
    return f.value, t.value


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _options(restart=30, maximum_iterations=300, relative_tolerance=1e-12, absolute_tolerance=0.0):
    return GmresOptions(restart, maximum_iterations, relative_tolerance, absolute_tolerance)


def gmres_settings() -> GmresOptions:
    """dielectric_gmres_settings (dielectric_fft.c:8)."""
    return GmresOptions.in_dll(lib(), "dielectric_gmres_settings")


def backend_set(name) -> bool:
    return bool(lib().vie_fft_backend_set(None if name is None else name.encode()))


def synchronize():
    return bool(lib().vie_hip_synchronize())


def copy_probe(nbytes, repeats=10):
    return lib().vie_hip_copy_probe(int(nbytes), int(repeats))


class FFTPlan:
    """vie_fft_plan (vie_fft.h) on the GPU."""

    def __init__(self, shape, xyz, pitch, k0, epsilon_r, byte_limit=1 << 62):
        L = lib()
        self.shape = np.ascontiguousarray(shape, dtype=np.uintp)
        self.xyz = np.ascontiguousarray(np.asarray(xyz).reshape(-1, 3), dtype=np.uintp)
        self.cells = len(self.xyz)
        p = ctypes.c_void_p()
        szp = ctypes.POINTER(ctypes.c_size_t)
        ok = L.vie_fft_create(ctypes.byref(p), self.shape.ctypes.data_as(szp), self.cells,
                              self.xyz.ctypes.data_as(szp), float(pitch), float(k0), CDouble.of(epsilon_r),
                              int(byte_limit))
        if not ok:
            raise ValueError(f"vie_fft_create rejected the plan: {last_error()}")
        self.p = p
        self.op = LinearOperator()
        L.vie_fft_operator(self.p, ctypes.byref(self.op))

    @property
    def order(self):
        return 3 * self.cells

    def set_permittivity(self, cell, er) -> bool:
        return bool(lib().vie_fft_set_permittivity(self.p, int(cell), CDouble.of(er)))

    def bytes(self):
        return lib().vie_fft_bytes(self.p)

    def set_permittivities(self, eps):
        e = np.ascontiguousarray(eps, dtype=np.complex128).reshape(-1)
        return bool(lib().vie_fft_set_permittivities(self.p, _ptr(e), len(e)))

    def backend(self):
        return lib().vie_fft_backend_name(self.p).decode()

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        f = ctypes.cast(self.op.apply, VectorMap)
        if not f(self.op.context, x.ctypes.data, y.ctypes.data):
            raise RuntimeError(f"apply failed: {last_error()}")
        return y

    def precondition(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        f = ctypes.cast(self.op.precondition, VectorMap)
        if not f(self.op.precondition_context, x.ctypes.data, y.ctypes.data):
            raise RuntimeError(f"precondition failed: {last_error()}")
        return y

    def solve(self, b, x=None, **opts):
        """vie_fft_solve: device-resident GMRES; returns (ok, x, report)."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b) if x is None else np.array(x, dtype=np.complex128)
        o = _options(**opts)
        rep = GmresReport()
        ok = lib().vie_fft_solve(self.p, _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(rep))
        return bool(ok), x, rep

    def gmres(self, b, x=None, **opts):
        """vie_gmres on this plan's operator (dispatches to the device solve)."""
        return gmres(self.op, b, x, **opts)

    def close(self):
        if getattr(self, "p", None):
            lib().vie_fft_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()


def gmres(op: LinearOperator, b, x=None, **opts):
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b) if x is None else np.array(x, dtype=np.complex128)
    o = _options(**opts)
    rep = GmresReport()
    ok = lib().vie_gmres(ctypes.byref(op), _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(rep))
    return bool(ok), x, rep


class CallbackOperator:
    """A vie_linear_operator whose callbacks are Python functions (tests)."""

    def __init__(self, order, apply, precondition=None):
        def wrap(fn):
            def cb(ctx, xin, yout):
                xi = np.ctypeslib.as_array((ctypes.c_double * (2 * order)).from_address(xin)).view(np.complex128)
                yo = np.ctypeslib.as_array((ctypes.c_double * (2 * order)).from_address(yout)).view(np.complex128)
                r = fn(xi.copy())
                if r is None:
                    return 0
                yo[:] = r
                return 1
            return VectorMap(cb)
        self._apply = wrap(apply)
        self._prec = wrap(precondition) if precondition else None
        self.op = LinearOperator(order, ctypes.cast(self._apply, ctypes.c_void_p), None,
                                 ctypes.cast(self._prec, ctypes.c_void_p) if self._prec else None, None)


@dataclass
class Cell:
    center: tuple
    size: tuple
    material_id: int


@dataclass
class MaterialSpec:
    """A material as the model holds it (dielectric.h:26-35): eps_r, a loss
    (positive imaginary-part magnitude) and a conductivity in S/m; or, with
    ``samples`` = ((frequency_mhz, eps_r, loss, conductivity_kind), ...) in
    ascending frequency, a tabulated material interpolated linearly in
    frequency (loss is a conductivity in S/m where conductivity_kind is 1)."""
    id: int
    epsilon_r: float
    epsilon_loss: float = 0.0
    conductivity: float = 0.0
    samples: tuple = ()


class Voxels:
    """A compact per-voxel resource (vie_hip_voxels_load): NEC2C-DIELECTRIC 4
    with one VOXELS record and a binary side file."""

    def __init__(self, path):
        self.v = ctypes.c_void_p()
        self.diag = Diagnostic()
        if not lib().vie_hip_voxels_load(ctypes.byref(self.v), str(path).encode(), ctypes.byref(self.diag)):
            self.v = None
            d = self.diag
            raise ValueError(f"{d.resource.decode()}:{d.line}: {d.record.decode()} {d.value.decode()}: "
                             f"{d.message.decode()}")
        shape = (ctypes.c_size_t * 3)()
        origin = (ctypes.c_double * 3)()
        pitch, cells = ctypes.c_double(), ctypes.c_size_t()
        lib().vie_hip_voxels_info(self.v, shape, ctypes.byref(pitch), origin, ctypes.byref(cells))
        self.shape, self.pitch, self.origin, self.cells = tuple(shape), pitch.value, tuple(origin), cells.value

    def cell_arrays(self):
        """(centers [cells, 3], sizes [cells, 3], material ids [cells])."""
        c = np.zeros((self.cells, 3))
        s = np.zeros((self.cells, 3))
        m = np.zeros(self.cells, dtype=np.int32)
        lib().vie_hip_voxels_cells(self.v, c.ctypes.data, s.ctypes.data, m.ctypes.data)
        return c, s, m

    def close(self):
        if getattr(self, "v", None):
            lib().vie_hip_voxels_free(self.v)
            self.v = None

    def __del__(self):
        self.close()


class System:
    """dielectric_fft_system (dielectric_fft.h) on the GPU."""

    def __init__(self, cells, materials, conductors, k0, pec=False, extra_bytes=0, pitch=None,
                 material_ids=None, ground=None):
        """cells: a list of Cell, or an (N, 3) array of cell centres together
        with ``pitch`` (and optional per-cell ``material_ids``, default 1).
        ground: the complex permittivity of a real ground below z = 0
        (dielectric_fft_prepare_ground), or None."""
        L = lib()
        self._mats = (Material * max(1, len(materials)))()
        self._samples = []
        for i, m in enumerate(materials):
            self._mats[i].id = m.id
            self._mats[i].epsilon_r = m.epsilon_r
            self._mats[i].epsilon_loss = m.epsilon_loss
            self._mats[i].conductivity = m.conductivity
            if m.samples:
                tab = (Sample * len(m.samples))()
                for j, (f, er, loss, kind) in enumerate(m.samples):
                    tab[j].material_id, tab[j].conductivity_kind = m.id, int(kind)
                    tab[j].frequency_mhz, tab[j].epsilon_r, tab[j].loss = float(f), float(er), float(loss)
                self._samples.append(tab)
                self._mats[i].tabulated = 1
                self._mats[i].samples = ctypes.cast(tab, ctypes.c_void_p)
                self._mats[i].sample_count = len(m.samples)
        self._cells = (MeshCell * max(1, len(cells)))()
        if isinstance(cells, np.ndarray):
            arr = np.ctypeslib.as_array(self._cells)
            arr["center"] = cells.reshape(-1, 3)
            arr["size"] = float(pitch)
            arr["material_id"] = 1 if material_ids is None else np.asarray(material_ids).reshape(-1)
        else:
            for i, c in enumerate(cells):
                self._cells[i].center[:] = list(c.center)
                self._cells[i].size[:] = list(c.size)
                self._cells[i].material_id = c.material_id
        self.model = Model()
        self.model.materials = self._mats
        self.model.material_count = len(materials)
        self.mesh = MeshResult()
        self.mesh.cells = self._cells
        self.mesh.cell_count = len(cells)
        self.diag = Diagnostic()
        s = ctypes.c_void_p()
        if ground is not None:
            g = complex(ground)
            ok = L.dielectric_fft_prepare_ground(ctypes.byref(s), ctypes.byref(self.model), ctypes.byref(self.mesh),
                                                 int(conductors), float(k0), int(extra_bytes), g.real, g.imag,
                                                 ctypes.byref(self.diag))
        else:
            fn = L.dielectric_fft_prepare_pec if pec else L.dielectric_fft_prepare
            ok = fn(ctypes.byref(s), ctypes.byref(self.model), ctypes.byref(self.mesh), int(conductors),
                    float(k0), int(extra_bytes), ctypes.byref(self.diag))
        self.message = self.diag.message.decode()
        if not ok:
            self.s = None
            raise ValueError(self.message)
        self.s = s
        self.c = int(conductors)
        self.v = 3 * len(cells)
        self.op = LinearOperator()

    @property
    def order(self):
        return self.c + self.v

    @classmethod
    def from_voxels(cls, voxels: "Voxels", conductors, k0, pec=False, extra_bytes=0):
        """dielectric_fft_prepare_voxels: the system of a compact voxel resource."""
        self = cls.__new__(cls)
        self.diag = Diagnostic()
        s = ctypes.c_void_p()
        ok = lib().dielectric_fft_prepare_voxels(ctypes.byref(s), voxels.v, int(conductors), float(k0),
                                                 int(extra_bytes), int(bool(pec)), ctypes.byref(self.diag))
        self.message = self.diag.message.decode()
        if not ok:
            self.s = None
            raise ValueError(self.message)
        self.s = s
        self.c = int(conductors)
        self.v = 3 * voxels.cells
        self.op = LinearOperator()
        return self

    def set(self, row, col, value) -> bool:
        p = lib().dielectric_fft_entry(self.s, int(row), int(col))
        if not p:
            return False
        arr = (ctypes.c_double * 2).from_address(p)
        arr[0], arr[1] = complex(value).real, complex(value).imag
        return True

    def set_blocks(self, C=None, B=None, D=None):
        """Write conductor blocks: C (c x c), B (c x v), D (v x c)."""
        c, v = self.c, self.v
        for r in range(c):
            if C is not None:
                for q in range(c):
                    self.set(r, q, C[r, q])
        # bulk writes straight into the entry storage
        if B is not None and c:
            base = lib().dielectric_fft_entry(self.s, 0, 0)
            rows = np.ctypeslib.as_array((ctypes.c_double * (2 * c * (c + v))).from_address(base))
            rows = rows.view(np.complex128).reshape(c, c + v)
            rows[:, c:] = B
        if D is not None and c and v:
            base = lib().dielectric_fft_entry(self.s, c, 0)
            dd = np.ctypeslib.as_array((ctypes.c_double * (2 * v * c)).from_address(base))
            dd.view(np.complex128).reshape(v, c)[:] = D
        lib().dielectric_fft_blocks_changed(self.s)

    def assemble_rows(self, k0, centers, sizes, wire_points=None, wire_tangents=None, patch_points=None,
                      patch_t1=None, patch_t2=None, patch_scale=None, ground=0):
        """dielectric_fft_assemble_conductor_rows: the B block on the device."""
        c, s, wp, wt, pp, t1, t2, ps = _conductor_arrays(centers, sizes, wire_points, wire_tangents, patch_points,
                                                         patch_t1, patch_t2, patch_scale)
        if len(c) != self.v // 3:
            raise ValueError("the system's cell count differs from the cells given")
        if not lib().dielectric_fft_assemble_conductor_rows(self.s, float(k0), c.ctypes.data, s.ctypes.data, len(wp),
                                                            wp.ctypes.data, wt.ctypes.data, len(pp), pp.ctypes.data,
                                                            t1.ctypes.data, t2.ctypes.data, ps.ctypes.data,
                                                            int(ground)):
            raise RuntimeError(f"conductor rows failed: {last_error()}")

    def assemble_columns(self, k0, wlam, centers, sizes, chi, segments, sources, patches=None, rkh=1.0, ground=0):
        """dielectric_fft_assemble_conductor_columns: the D block on the device."""
        c, s, ch, seg, n, pat, m, src = _columns_arrays(centers, sizes, chi, segments, patches, sources)
        if len(c) != self.v // 3 or len(src) != self.c:
            raise ValueError("the system's cell count differs from the cells given, or not one column per conductor")
        if not lib().dielectric_fft_assemble_conductor_columns(self.s, float(k0), float(wlam), c.ctypes.data,
                                                               s.ctypes.data, ch.ctypes.data, n,
                                                               *[a.ctypes.data for a in seg], m,
                                                               *[a.ctypes.data for a in pat], src.ctypes.data,
                                                               float(rkh), int(ground)):
            raise RuntimeError(f"conductor columns failed: {last_error()}")

    def set_permittivities(self, eps):
        """Per-cell permittivity in the caller's cell order (vie_fft_set_permittivities)."""
        e = np.ascontiguousarray(eps, dtype=np.complex128).reshape(-1)
        plan = lib().dielectric_fft_plan(self.s)
        if not plan or not lib().vie_fft_set_permittivities(plan, _ptr(e), len(e)):
            raise ValueError("permittivity rejected")

    def operator(self) -> LinearOperator:
        lib().dielectric_fft_get_operator(self.s, ctypes.byref(self.op))
        return self.op

    def storage(self):
        return lib().dielectric_fft_storage(self.s)

    def backend(self):
        return lib().dielectric_fft_backend(self.s).decode()

    def apply(self, x):
        op = self.operator()
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        if not ctypes.cast(op.apply, VectorMap)(op.context, x.ctypes.data, y.ctypes.data):
            raise RuntimeError(f"apply failed: {last_error()}")
        return y

    def apply_ex(self, x, precision=FP64, jacobi=False):
        """y = A M^{-1} x (jacobi) or A x at an explicit device precision."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        if not lib().dielectric_fft_apply_ex(self.s, int(precision), int(bool(jacobi)), _ptr(x), _ptr(y)):
            raise RuntimeError(f"apply failed: {last_error()}")
        return y

    def precondition(self, x):
        op = self.operator()
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.order, dtype=np.complex128)
        if not ctypes.cast(op.precondition, VectorMap)(op.precondition_context, x.ctypes.data, y.ctypes.data):
            raise RuntimeError(f"precondition failed: {last_error()}")
        return y

    def solve(self, b, x=None, precision=None, inner_reduction=0.0, recycle=0, **opts):
        """dielectric_fft_solve; precision None keeps the library default."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b) if x is None else np.array(x, dtype=np.complex128)
        o = _options(**opts)
        rep = GmresReport()
        L = lib()
        if precision is not None:
            old = SolveOptions()
            L.vie_hip_get_solve_options(ctypes.byref(old))
            L.vie_hip_set_solve_options(ctypes.byref(SolveOptions(precision, inner_reduction, recycle)))
        try:
            ok = L.dielectric_fft_solve(self.s, _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(rep))
        finally:
            if precision is not None:
                L.vie_hip_set_solve_options(ctypes.byref(old))
        return bool(ok), x, rep

    def gmres(self, b, x=None, **opts):
        return gmres(self.operator(), b, x, **opts)

    # resident work
    def upload_rhs(self, b):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        if not lib().dielectric_fft_upload_rhs(self.s, _ptr(b)):
            raise RuntimeError(last_error())

    def reserve(self, **opts):
        """dielectric_fft_reserve: make a later resident solve allocation-free."""
        o = _options(**opts)
        if not lib().dielectric_fft_reserve(self.s, ctypes.byref(o)):
            raise RuntimeError(last_error())

    def solve_resident(self, precision=FP64, inner_reduction=0.0, recycle=0, **opts):
        o = _options(**opts)
        so = SolveOptions(precision, inner_reduction, recycle)
        rep = GmresReport()
        ok = lib().dielectric_fft_solve_resident(self.s, ctypes.byref(o), ctypes.byref(so), ctypes.byref(rep))
        return bool(ok), rep

    def download_solution(self):
        x = np.zeros(self.order, dtype=np.complex128)
        if not lib().dielectric_fft_download_solution(self.s, _ptr(x)):
            raise RuntimeError(last_error())
        return x

    def tune(self):
        """vie_fft_tune on the system's plan: the chosen launch variant of each
        pass, {"fp64": [5], "fp32": [5]} (None without cells)."""
        plan = lib().dielectric_fft_plan(self.s)
        if not plan:
            return None
        v = (ctypes.c_int * 10)()
        if not lib().vie_fft_tune(plan, v):
            raise RuntimeError(f"tune failed: {last_error()}")
        return {"fp64": list(v[:5]), "fp32": list(v[5:])}

    def set_variants(self, fp64, fp32):
        """vie_fft_set_variants: five pass variants per precision; False when one does not exist."""
        plan = lib().dielectric_fft_plan(self.s)
        v = (ctypes.c_int * 10)(*(list(fp64) + list(fp32)))
        return bool(plan) and bool(lib().vie_fft_set_variants(plan, v))

    def time_apply(self, precision=FP64, repeats=20):
        ms = ctypes.c_double()
        passes = (ctypes.c_double * 6)()
        if not lib().dielectric_fft_time_apply(self.s, precision, repeats, ctypes.byref(ms), passes):
            raise RuntimeError(last_error())
        return ms.value, list(passes)

    @property
    def escalated(self):
        """True when the last MIXED device solve finished in FP64 cycles."""
        return bool(lib().dielectric_fft_last_solve_escalated(self.s))

    @property
    def recycled(self):
        """Recycled-subspace size the last device solve built and used (0: plain cycles)."""
        return int(lib().dielectric_fft_last_solve_recycled(self.s))

    def solve_stats(self, reset=False):
        """(mean basis column, matvec Arnoldi steps) of the resident solves since the last reset."""
        mj, st = ctypes.c_double(), ctypes.c_ulonglong()
        lib().dielectric_fft_solve_stats(self.s, ctypes.byref(mj), ctypes.byref(st), int(reset))
        return mj.value, st.value

    def time_arnoldi(self, precision=FP64, j=60, repeats=20):
        """(ms, bytes) of the delayed-CGS2 reduction and fused update at column j."""
        v = [ctypes.c_double() for _ in range(4)]
        if not lib().dielectric_fft_time_arnoldi(self.s, precision, j, repeats, *[ctypes.byref(x) for x in v]):
            raise RuntimeError(f"time_arnoldi failed: {last_error()}")
        return {"ms_reduce": v[0].value, "ms_update": v[1].value, "bytes_reduce": v[2].value,
                "bytes_update": v[3].value}

    def apply_bytes(self, precision=FP64):
        passes = (ctypes.c_double * 6)()
        tot = ctypes.c_double()
        lib().dielectric_fft_apply_bytes(self.s, precision, passes, ctypes.byref(tot))
        return tot.value, list(passes)

    def close(self):
        if getattr(self, "s", None):
            lib().dielectric_fft_free(self.s)
            self.s = None

    def __del__(self):
        self.close()


class Batch:
    """Many device-resident solves in flight (vie_hip_batch_*).

    Jobs are submitted per (system, right-hand side); the library interleaves
    their Arnoldi steps and overlaps jobs of different systems on the GPU.
    """

    def __init__(self, share_subspaces=False):
        self.b = lib().vie_hip_batch_create()
        if not self.b:
            raise RuntimeError(f"batch create failed: {last_error()}")
        self._orders = {}
        if share_subspaces:
            self.share_subspaces(True)

    def share_subspaces(self, on=True):
        """vie_hip_batch_share_subspaces: jobs of one system submitted from now
        on share the recycled subspace of the first of them to finish a cycle."""
        lib().vie_hip_batch_share_subspaces(self.b, 1 if on else 0)

    def shared_stats(self):
        """(subspaces published, subspaces adopted) by the finished jobs."""
        p, a = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        lib().vie_hip_batch_shared_stats(self.b, ctypes.byref(p), ctypes.byref(a))
        return p.value, a.value

    @staticmethod
    def job_bytes(system, precision=FP64, recycle=0, **opts):
        o = _options(**opts)
        so = SolveOptions(precision, 0.0, recycle)
        return lib().vie_hip_batch_job_bytes(system.s, ctypes.byref(o), ctypes.byref(so))

    def reserve(self, count, system, **opts):
        o = _options(**opts)
        if not lib().vie_hip_batch_reserve(self.b, count, system.s, ctypes.byref(o)):
            raise RuntimeError(f"batch reserve failed: {last_error()}")

    def submit(self, system, rhs, precision=FP64, inner_reduction=0.0, recycle=0, **opts):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        if rhs.size != system.order:
            raise ValueError(f"right-hand side has {rhs.size} entries, the system {system.order}")
        o = _options(**opts)
        so = SolveOptions(precision, inner_reduction, recycle)
        jid = lib().vie_hip_batch_submit(self.b, system.s, _ptr(rhs), ctypes.byref(o), ctypes.byref(so))
        if jid <= 0:
            raise RuntimeError(f"batch submit failed: {last_error()}")
        self._orders[jid] = system.order
        return jid

    def progress(self, max_ready=1024):
        ids = (ctypes.c_longlong * max_ready)()
        running = ctypes.c_size_t()
        k = lib().vie_hip_batch_progress(self.b, ids, max_ready, ctypes.byref(running))
        return [ids[q] for q in range(k)], running.value

    @property
    def running(self):
        return lib().vie_hip_batch_running(self.b)

    def wait(self, jid):
        if not lib().vie_hip_batch_wait(self.b, jid):
            raise RuntimeError(last_error())

    def wait_all(self):
        lib().vie_hip_batch_wait_all(self.b)

    def result(self, jid, release=True):
        """(converged, x, report, seconds from submission to completion)."""
        x = np.zeros(self._orders[jid], dtype=np.complex128)
        rep = GmresReport()
        sec = ctypes.c_double()
        ok = lib().vie_hip_batch_result(self.b, jid, _ptr(x), ctypes.byref(rep), ctypes.byref(sec))
        if release:
            lib().vie_hip_batch_release(self.b, jid)
            self._orders.pop(jid, None)
        return bool(ok), x, rep, sec.value

    def group_stats(self):
        """(multi-job steps issued, job steps they covered): one launch per
        kernel served every member of a multi-job step."""
        g, j = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        lib().vie_hip_batch_group_stats(self.b, ctypes.byref(g), ctypes.byref(j))
        return g.value, j.value

    def close(self):
        if getattr(self, "b", None):
            lib().vie_hip_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        self.close()


def lattice_cells(shape, pitch, origin=(0.0, 0.0, 0.0), material_ids=None, mask=None):
    """Cells of a box lattice (x slowest, z fastest), optionally masked."""
    nx, ny, nz = shape
    out = []
    idx = 0
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                if mask is not None and not mask[x, y, z]:
                    continue
                mid = 1 if material_ids is None else int(material_ids[x, y, z])
                out.append(Cell((origin[0] + pitch * x, origin[1] + pitch * y, origin[2] + pitch * z),
                                (pitch, pitch, pitch), mid))
                idx += 1
    return out
