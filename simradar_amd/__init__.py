"""simradar_amd -- MI355X-native per-pulse scatterer engine behind SimRadar's RS_* C API.

The product is the shared library ``libsimradar_amd.so`` built from ``csrc/``
(plain-C RS_* host runtime + HIP device layer for gfx950).  This module only
binds that library with ctypes (for tests, the benchmark and Python callers);
every computation runs in the library.  Importing the bindings never falls
back to anything else: a missing library raises ImportError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimradar_amd.so")

RS_MAX_DEBRIS_TYPES = 8

# rs_enum.h concept bits
CONCEPT_NULL = 0
CONCEPT_CROSS_POL = 1
CONCEPT_DRAGGED = 1 << 1
CONCEPT_TRANSPARENT = 1 << 2
CONCEPT_BOUNDED = 1 << 3
CONCEPT_UNIFORM_DSD = 1 << 4
CONCEPT_FIXED = 1 << 5
CONCEPT_VERTICAL = 1 << 6
CONCEPT_DFF = 1 << 7

STATUS_POPULATION_DEFINED = 1 << 1
STATUS_ORIGINS = 1 << 2
STATUS_WORKERS = 1 << 3
STATUS_POPULATED = 1 << 4
STATUS_SIGNAL_UPDATE = 1 << 5
STATUS_DEBRIS_RCS_UPDATE = 1 << 6

PHASE_F64 = 0
PHASE_F32 = 1

FOLD_LANE = 0       # gate-sum strategies of rsd_set_fold (include/rsd.h)
FOLD_WAVE64 = 1
FOLD_WAVE32 = 2
FOLD_WINDOW = 3

OBJ_LEAF = 1
OBJ_LEAF_BIG = 2
OBJ_WOODBOARD_2X4 = 3
OBJ_WOODBOARD_4X8 = 4
OBJ_METAL_SHEET = 5
OBJ_BRICK = 6

STREAMS = {"pos": 0, "vel": 1, "ori": 2, "tum": 3, "aux": 4, "rcs": 5, "sig": 6, "rnd": 7}

EXPORTED_RS = [
    "RS_gpu_count", "RS_version_string", "RS_indent_copy", "RS_init_with_path", "RS_init_for_selected_gpu",
    "RS_init_for_cpu_verbose", "RS_init_verbose", "RS_init", "RS_free", "RS_make_pulse_params",
    "RS_set_concept", "RS_simulation_concept_string", "RS_simulation_concept_bulleted_string", "RS_set_prt",
    "RS_set_lambda", "RS_set_density", "RS_set_antenna_params", "RS_set_tx_params", "RS_set_sampling_spacing",
    "RS_set_scan_box", "RS_set_scan_extent", "RS_set_beam_pos", "RS_set_verbosity", "RS_set_debris_count",
    "RS_get_debris_count", "RS_get_worker_debris_count", "RS_get_all_worker_debris_counts", "RS_get_domain",
    "RS_set_dsd", "RS_set_dsd_to_mp", "RS_set_dsd_to_mp_with_sizes", "RS_set_range_weight",
    "RS_set_range_weight_to_triangle", "RS_set_angular_weight", "RS_set_angular_weight_to_standard",
    "RS_set_angular_weight_to_double_cone", "RS_set_vel_data_to_config", "RS_set_vel_data_to_LES_table",
    "RS_set_vel_data_to_uniform", "RS_set_vel_data_to_cube27", "RS_set_vel_data_to_cube125", "RS_clear_vel_data",
    "RS_set_debris_flux_field_by_pdf", "RS_set_debris_flux_field_by_icdf",
    "RS_set_debris_flux_field_to_center_cell_of_3x3", "RS_set_debris_flux_field_to_checker_board",
    "RS_set_debris_flux_field_to_checker_board_stretched", "RS_set_debris_flux_field_from_LES",
    "RS_set_scan_pattern", "RS_set_scan_pattern_with_string", "RS_set_random_seed", "RS_add_debris",
    "RS_io_test", "RS_populate", "RS_upload", "RS_download", "RS_download_position_only",
    "RS_download_orientation_only", "RS_download_pulse_only", "RS_compute_rcs_ellipsoids", "RS_advance_time",
    "RS_advance_beam", "RS_make_pulse", "RS_table_init", "RS_table_free", "RS_table2d_init", "RS_table2d_free",
    "RS_table3d_init", "RS_table3d_free", "RS_show_radar_params", "RS_show_scat_pos", "RS_show_scat_sig",
    "RS_show_scat_att", "RS_show_pulse", "RS_suggest_scan_domain", "RS_revise_debris_counts_to_gpu_preference",
    "RS_simulation_description", "RS_set_vel_data", "RS_set_adm_data", "RS_set_rcs_data",
    "RS_set_adm_data_to_ADM_table", "RS_set_adm_data_to_unity", "RS_clear_adm_data",
    "RS_set_rcs_data_to_RCS_table", "RS_set_rcs_data_to_unity", "RS_clear_rcs_data",
    "RS_set_rcs_ellipsoid_table", "RS_revise_population", "RS_update_origins_offsets", "RS_merge_pulse_tmp",
    "RS_init_host_only", "RS_set_population", "RS_set_phase_precision", "RS_set_les_time_interpolation", "RS_get_pulse_f64",
    "RS_set_moment_pulses", "RS_get_moments", "RS_set_range_cap", "RS_batch_round",
    "RS_get_descriptor", "RS_radar_gain", "RS_engine", "RS_get_info", "RS_get_host_stream", "RS_get_params",
    "RS_get_dsd", "RS_suggest_scan_domain_values", "RS_set_scan_box_values", "RS_get_les_frame", "RS_get_les_cpxx", "RS_get_debris_table", "RS_get_debris_flux_field",
]
EXPORTED_RSD = [
    "rsd_device_count", "rsd_device_props", "rsd_open", "rsd_close", "rsd_last_error", "rsd_alloc_state",
    "rsd_upload_state", "rsd_download_state", "rsd_set_wind", "rsd_set_cpxx", "rsd_set_adm", "rsd_set_rcs", "rsd_set_drop_rcs", "rsd_set_dff", "rsd_set_dff_slot",
    "rsd_set_range_weight", "rsd_set_angular_weight", "rsd_advance", "rsd_make_pulse", "rsd_make_pulse_noalloc", "rsd_read_pulse",
    "rsd_sync", "rsd_timer_start", "rsd_timer_stop", "rsd_set_launch", "rsd_math_probe", "rsd_profile",
    "rsd_profile_read", "rsd_io_test", "rsd_set_sort", "rsd_get_sort", "rsd_set_fold", "rsd_set_iq_ring", "rsd_moments",
    "rsd_set_position_precision", "rsd_download_pos_lo", "rsd_batch_create", "rsd_batch_destroy",
    "rsd_batch_launch", "rsd_batch_sync", "rsd_batch_set_launch", "rsd_batch_get_launch", "rsd_batch_set_graph",
    "rsd_batch_timer_start", "rsd_batch_timer_stop", "rsd_batch_last_error",
    "rsd_mlp_create", "rsd_mlp_destroy", "rsd_mlp_set_params", "rsd_mlp_get_params", "rsd_mlp_set_output_scaling",
    "rsd_mlp_forward", "rsd_mlp_gradient", "rsd_mlp_set_data", "rsd_mlp_train", "rsd_mlp_sync", "rsd_mlp_last_error",
    "rsd_debris_samples", "rsd_debris_state_samples", "rsd_set_debris_surrogate", "rsd_wind_index_table",
]


class RSParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "c", "prt", "loss", "lambda_", "tx_power_watt", "antenna_gain_dbi", "antenna_bw_deg", "tau",
        "range_start", "range_end", "range_delta", "azimuth_start_deg", "azimuth_end_deg", "azimuth_delta_deg",
        "elevation_start_deg", "elevation_end_deg", "elevation_delta_deg", "domain_pad_factor", "body_per_cell",
        "prf", "va", "fn", "antenna_bw_rad", "dr")] + [("range_count", ctypes.c_uint32)]


class RSInfo(ctypes.Structure):
    _fields_ = [("num_scats", ctypes.c_uint64), ("num_types", ctypes.c_uint64),
                ("counts", ctypes.c_uint64 * RS_MAX_DEBRIS_TYPES), ("origins", ctypes.c_uint64 * RS_MAX_DEBRIS_TYPES),
                ("sim_tic", ctypes.c_float), ("sim_toc", ctypes.c_float), ("status", ctypes.c_uint32),
                ("sim_concept", ctypes.c_uint32), ("random_seed", ctypes.c_uint32), ("vel_idx", ctypes.c_uint32),
                ("les_id", ctypes.c_uint32), ("range_count", ctypes.c_uint32), ("last_error", ctypes.c_int32),
                ("dsd_count", ctypes.c_int32), ("adm_count", ctypes.c_uint32), ("rcs_count", ctypes.c_uint32),
                ("les_blend", ctypes.c_uint32), ("wind_alpha", ctypes.c_float)]


class RSPolar(ctypes.Structure):
    _fields_ = [("r", ctypes.c_float), ("a", ctypes.c_float), ("e", ctypes.c_float), ("w", ctypes.c_float)]


class RSBox(ctypes.Structure):
    _fields_ = [("origin", RSPolar), ("size", RSPolar)]


class CLFloat4(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float), ("w", ctypes.c_float)]


class RSVolume(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("ox", "oy", "oz", "sx", "sy", "sz")]


class RSMakePulseParams(ctypes.Structure):
    _fields_ = [("num_scats", ctypes.c_uint), ("user_max_groups", ctypes.c_uint),
                ("user_max_work_items", ctypes.c_uint), ("cl_pass_2_method", ctypes.c_uint),
                ("range_count", ctypes.c_uint), ("range_start", ctypes.c_float), ("range_delta", ctypes.c_float),
                ("entry_counts", ctypes.c_uint * 2), ("group_counts", ctypes.c_uint * 2),
                ("global_", ctypes.c_size_t * 2), ("local", ctypes.c_size_t * 2),
                ("local_mem_size", ctypes.c_size_t * 2)]


class RSDStep(ctypes.Structure):
    _fields_ = [("sim_desc", ctypes.c_float * 16), ("wind_slot", ctypes.c_int32),
                ("adm_index", ctypes.c_int32 * 8), ("rcs_index", ctypes.c_int32 * 8),
                ("background_kind", ctypes.c_uint32), ("wind_blend", ctypes.c_uint32), ("wind_alpha", ctypes.c_float)]


class RSDPulse(ctypes.Structure):
    _fields_ = [("sim_desc", ctypes.c_float * 16), ("two_over_lambda", ctypes.c_double),
                ("range_start", ctypes.c_float), ("range_delta", ctypes.c_float), ("range_count", ctypes.c_uint32),
                ("rcs_index", ctypes.c_int32 * 8), ("debris_rcs_refresh", ctypes.c_uint32),
                ("phase_mode", ctypes.c_uint32)]


_lib = None


def _declare(lib):
    P, F, D, U, I = ctypes.c_void_p, ctypes.c_float, ctypes.c_double, ctypes.c_uint, ctypes.c_int
    S = ctypes.c_size_t
    sig = {
        "RS_gpu_count": (U, []),
        "RS_version_string": (ctypes.c_char_p, []),
        "RS_init_with_path": (P, [ctypes.c_char_p, ctypes.c_char, ctypes.c_uint8, ctypes.c_ssize_t, ctypes.c_char]),
        "RS_init_verbose": (P, [ctypes.c_char]),
        "RS_init": (P, []),
        "RS_init_host_only": (P, [ctypes.c_char]),
        "RS_init_for_cpu_verbose": (P, [ctypes.c_char]),
        "RS_free": (None, [P]),
        "RS_make_pulse_params": (RSMakePulseParams, [U, U, U, U, F, F, U]),
        "RS_set_concept": (None, [P, U]),
        "RS_set_debris_flux_field_by_pdf": (None, [P, P, P]),
        "RS_set_debris_flux_field_by_icdf": (None, [P, P, P]),
        "RS_set_debris_flux_field_to_center_cell_of_3x3": (None, [P]),
        "RS_set_debris_flux_field_to_checker_board": (None, [P, I]),
        "RS_set_debris_flux_field_to_checker_board_stretched": (None, [P, P]),
        "RS_set_debris_flux_field_from_LES": (None, [P, P]),
        "RS_simulation_concept_string": (ctypes.c_char_p, [P]),
        "RS_set_prt": (None, [P, F]),
        "RS_set_lambda": (None, [P, F]),
        "RS_set_density": (None, [P, F]),
        "RS_set_antenna_params": (None, [P, F, F]),
        "RS_set_tx_params": (None, [P, F, F]),
        "RS_set_sampling_spacing": (None, [P, F, F, F]),
        "RS_set_scan_extent": (None, [P, F, F, F, F, F, F, F, F, F]),
        "RS_set_beam_pos": (None, [P, F, F]),
        "RS_set_verbosity": (None, [P, ctypes.c_char]),
        "RS_set_debris_count": (None, [P, I, S]),
        "RS_get_debris_count": (S, [P, I]),
        "RS_get_domain": (RSVolume, [P]),
        "RS_set_dsd": (None, [P, P, P, I, ctypes.c_char]),
        "RS_set_dsd_to_mp": (None, [P]),
        "RS_set_range_weight": (None, [P, P, F, F, U]),
        "RS_set_range_weight_to_triangle": (None, [P, F]),
        "RS_set_angular_weight": (None, [P, P, F, F, U]),
        "RS_set_angular_weight_to_standard": (None, [P, F]),
        "RS_set_angular_weight_to_double_cone": (None, [P, F]),
        "RS_set_vel_data_to_config": (None, [P, ctypes.c_char_p]),
        "RS_set_vel_data_to_uniform": (None, [P, CLFloat4]),
        "RS_set_vel_data_to_cube27": (None, [P]),
        "RS_set_vel_data_to_cube125": (None, [P]),
        "RS_set_scan_pattern_with_string": (None, [P, ctypes.c_char_p]),
        "RS_set_random_seed": (None, [P, U]),
        "RS_add_debris": (None, [P, I, S]),
        "RS_populate": (None, [P]),
        "RS_upload": (None, [P]),
        "RS_download": (None, [P]),
        "RS_download_position_only": (None, [P]),
        "RS_download_orientation_only": (None, [P]),
        "RS_download_pulse_only": (None, [P]),
        "RS_advance_time": (None, [P]),
        "RS_advance_beam": (None, [P]),
        "RS_make_pulse": (None, [P]),
        "RS_revise_debris_counts_to_gpu_preference": (None, [P]),
        "RS_simulation_description": (ctypes.c_char_p, [P]),
        "RS_set_population": (None, [P, S]),
        "RS_set_les_time_interpolation": (None, [P, I]),
        "RS_set_phase_precision": (None, [P, ctypes.c_uint32]),
        "RS_get_pulse_f64": (I, [P, P]),
        "RS_set_moment_pulses": (None, [P, I]),
        "RS_set_range_cap": (None, [P, F]),
        "RS_batch_round": (I, [P, I, P]),
        "RS_get_moments": (I, [P, I, P]),
        "RS_get_descriptor": (I, [P, I, P]),
        "RS_radar_gain": (F, [P]),
        "RS_engine": (P, [P]),
        "RS_get_info": (None, [P, ctypes.POINTER(RSInfo)]),
        "RS_get_host_stream": (P, [P, I]),
        "RS_get_params": (None, [P, ctypes.POINTER(RSParams)]),
        "RS_get_dsd": (I, [P, P, P, P, I]),
        "RS_get_debris_flux_field": (I, [P, I, P, I, P]),
        "RS_suggest_scan_domain_values": (None, [P, P]),
        "RS_get_les_frame": (P, [P, I, P]),
        "RS_get_les_cpxx": (P, [P, I, P]),
        "RS_get_debris_table": (I, [P, I, I, P, P]),
        "RS_set_scan_box_values": (None, [P, P]),
        "rsd_device_count": (I, []),
        "rsd_device_props": (I, [I, P, P, P]),
        "rsd_open": (I, [I, P]),
        "rsd_close": (None, [P]),
        "rsd_last_error": (ctypes.c_char_p, [P]),
        "rsd_alloc_state": (I, [P, S, P, P]),
        "rsd_upload_state": (I, [P, P, P, P, P, P, P, P]),
        "rsd_download_state": (I, [P, P, P, P, P, P, P, P, P]),
        "rsd_set_wind": (I, [P, I, P, I, I, I, P]),
        "rsd_set_cpxx": (I, [P, I, P, I, I, I]),
        "rsd_set_adm": (I, [P, I, P, P, I, I, P]),
        "rsd_set_rcs": (I, [P, I, P, P, I, I, P]),
        "rsd_set_drop_rcs": (I, [P, P, I, P]),
        "rsd_set_dff": (I, [P, P, I, P]),
        "rsd_set_dff_slot": (I, [P, I, P, I, P]),
        "rsd_set_range_weight": (I, [P, P, I, P]),
        "rsd_set_angular_weight": (I, [P, P, I, P]),
        "rsd_advance": (I, [P, ctypes.POINTER(RSDStep)]),
        "rsd_make_pulse": (I, [P, ctypes.POINTER(RSDPulse)]),
        "rsd_make_pulse_noalloc": (I, [P, ctypes.POINTER(RSDPulse), P]),
        "rsd_read_pulse": (I, [P, P, ctypes.c_uint32]),
        "rsd_sync": (I, [P]),
        "rsd_timer_start": (I, [P]),
        "rsd_timer_stop": (I, [P, ctypes.POINTER(ctypes.c_double)]),
        "rsd_set_launch": (I, [P, I, I]),
        "rsd_set_sort": (I, [P, I]),
        "rsd_get_sort": (I, [P]),
        "rsd_set_fold": (I, [P, I]),
        "rsd_set_iq_ring": (I, [P, I, I]),
        "rsd_moments": (I, [P, I, P, P]),
        "rsd_set_position_precision": (I, [P, I]),
        "rsd_download_pos_lo": (I, [P, P]),
        "rsd_batch_create": (I, [P, I, P]),
        "rsd_batch_destroy": (None, [P]),
        "rsd_batch_launch": (I, [P]),
        "rsd_batch_sync": (I, [P]),
        "rsd_batch_set_launch": (I, [P, I, I]),
        "rsd_batch_get_launch": (I, [P, P, P, P]),
        "rsd_batch_set_graph": (I, [P, I]),
        "rsd_batch_timer_start": (I, [P]),
        "rsd_batch_timer_stop": (I, [P, ctypes.POINTER(ctypes.c_double)]),
        "rsd_batch_last_error": (ctypes.c_char_p, [P]),
        "rsd_mlp_create": (I, [I, ctypes.c_uint64, P]),
        "rsd_mlp_destroy": (None, [P]),
        "rsd_mlp_set_params": (I, [P, P]),
        "rsd_mlp_get_params": (I, [P, P]),
        "rsd_mlp_set_output_scaling": (I, [P, P, P]),
        "rsd_mlp_forward": (I, [P, P, P, S]),
        "rsd_mlp_gradient": (I, [P, P, P, S, P, ctypes.POINTER(ctypes.c_double)]),
        "rsd_mlp_set_data": (I, [P, P, P, S]),
        "rsd_mlp_train": (I, [P, I, I, F, ctypes.POINTER(ctypes.c_double)]),
        "rsd_mlp_sync": (I, [P]),
        "rsd_mlp_last_error": (ctypes.c_char_p, [P]),
        "rsd_debris_samples": (I, [P, I, I, S, ctypes.c_uint64, P, P]),
        "rsd_debris_state_samples": (I, [P, I, I, I, I, P, S, P, P, P]),
        "rsd_set_debris_surrogate": (I, [P, I, P]),
        "rsd_math_probe": (I, [I, I, P, P, P, S]),
        "rsd_wind_index_table": (I, [P, P, P, P, P]),
        "rsd_profile": (I, [P, I]),
        "rsd_io_test": (I, [P, S, I, ctypes.POINTER(ctypes.c_double)]),
        "rsd_profile_read": (I, [P, I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load():
    """Load the engine library (built by ``python -m simradar_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("simradar_amd engine library not built: %s (run python -m simradar_amd.build)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(lib)
        _lib = lib
    return _lib


def gpu_count():
    return int(load().rsd_device_count())


def wind_index_table(les_desc, sim_desc):
    """The engine's stretched-grid index tables for an LES descriptor and a box
    (rsd_wind_index_table; host only): rows[768, 4] (x, y, z: 256 cubic pieces
    each), ih[3] pieces per metre, jm[3] last piece index."""
    les = np.ascontiguousarray(les_desc, np.float32)
    sd = np.ascontiguousarray(sim_desc, np.float32)
    rows = np.zeros((768, 4), np.float32)
    ih = np.zeros(3, np.float32)
    jm = np.zeros(3, np.float32)
    rc = load().rsd_wind_index_table(_ptr(les), _ptr(sd), _ptr(rows), _ptr(ih), _ptr(jm))
    if rc:
        raise RuntimeError("rsd_wind_index_table failed (%d)" % rc)
    return rows, ih, jm


def math_probe(fn, a, b=None, device=0):
    """Evaluate device function `fn` (see rsd_math_probe) on float32 inputs."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = a if b is None else np.ascontiguousarray(b, np.float32).ravel()
    out = np.empty_like(a)
    rc = load().rsd_math_probe(device, fn, _ptr(a), _ptr(b), _ptr(out), a.size)
    if rc:
        raise RuntimeError("rsd_math_probe failed (%d)" % rc)
    return out


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Radar:
    """Python mirror of the RS_* interface (one RSHandle).  Method names follow
    the C API without the RS_ prefix; arrays come back as numpy views/copies."""

    def __init__(self, verb=0, host_only=False):
        self.lib = load()
        self.H = self.lib.RS_init_host_only(verb) if host_only else self.lib.RS_init_verbose(verb)
        if not self.H:
            raise RuntimeError("RS_init returned NULL (no GPU device)" if not host_only else "RS_init_host_only failed")
        self.host_only = host_only

    def free(self):
        if self.H:
            self.lib.RS_free(self.H)
            self.H = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __getattr__(self, name):
        fn = getattr(load(), "RS_" + name, None)
        if fn is None:
            raise AttributeError(name)
        H = self.__dict__.get("H")
        return lambda *a: fn(H, *a)

    # setters with argument conversion
    def suggest_scan_domain(self):
        """Box as (origin r, a, e, w, size r, a, e, w)."""
        out = np.zeros(8, np.float32)
        self.lib.RS_suggest_scan_domain_values(self.H, _ptr(out))
        return out

    def set_scan_box(self, box8):
        box8 = np.ascontiguousarray(box8, np.float32)
        self.lib.RS_set_scan_box_values(self.H, _ptr(box8))

    def set_dsd(self, nd, diameters, name=3):
        nd = np.ascontiguousarray(nd, np.float32)
        d = np.ascontiguousarray(diameters, np.float32)
        self.lib.RS_set_dsd(self.H, _ptr(nd), _ptr(d), len(nd), name)

    def set_range_weight(self, w, start, delta):
        w = np.ascontiguousarray(w, np.float32)
        self.lib.RS_set_range_weight(self.H, _ptr(w), start, delta, len(w))

    def set_angular_weight(self, w, start, delta):
        w = np.ascontiguousarray(w, np.float32)
        self.lib.RS_set_angular_weight(self.H, _ptr(w), start, delta, len(w))

    def set_vel_data_to_config(self, name="suctvort"):
        self.lib.RS_set_vel_data_to_config(self.H, name.encode())

    def set_vel_data_to_uniform(self, u, v, w):
        self.lib.RS_set_vel_data_to_uniform(self.H, CLFloat4(u, v, w, 0.0))

    def set_scan_pattern(self, s):
        self.lib.RS_set_scan_pattern_with_string(self.H, s.encode())

    # accessors
    def info(self):
        out = RSInfo()
        self.lib.RS_get_info(self.H, ctypes.byref(out))
        return out

    def params(self):
        out = RSParams()
        self.lib.RS_get_params(self.H, ctypes.byref(out))
        return out

    def descriptor(self, which):
        out = np.zeros(16, np.float32)
        n = self.lib.RS_get_descriptor(self.H, which, _ptr(out))
        return out[:n].copy()

    def sim_desc(self):
        return self.descriptor(0)

    def stream(self, name, copy=True):
        inf = self.info()
        n = int(inf.num_scats)
        p = self.lib.RS_get_host_stream(self.H, STREAMS[name])
        if not p:
            return None
        dt = np.uint32 if name == "rnd" else np.float32
        buf = (ctypes.c_char * (n * 16)).from_address(p)
        a = np.frombuffer(buf, dtype=dt).reshape(n, 4)
        return a.copy() if copy else a

    def set_stream(self, name, values):
        dst = self.stream(name, copy=False)
        dst[...] = values

    def pulse(self):
        R = int(self.info().range_count)
        p = self.lib.RS_get_host_stream(self.H, 8)
        buf = (ctypes.c_char * (R * 16)).from_address(p)
        return np.frombuffer(buf, dtype=np.float32).reshape(R, 4).copy()

    def pulse_f64(self):
        R = int(self.info().range_count)
        out = np.zeros((R, 4), np.float64)
        n = self.lib.RS_get_pulse_f64(self.H, _ptr(out))
        if n != R:
            raise RuntimeError("RS_get_pulse_f64 failed")
        return out

    def les_frame(self, frame):
        dims = np.zeros(3, np.uint32)
        p = self.lib.RS_get_les_frame(self.H, frame, _ptr(dims))
        if not p:
            return None
        nx, ny, nz = (int(v) for v in dims)
        buf = (ctypes.c_char * (nx * ny * nz * 16)).from_address(p)
        return np.frombuffer(buf, dtype=np.float32).reshape(nz, ny, nx, 4).copy()

    def les_cpxx(self, frame):
        """(nz, ny, nx, 4) float32 (cn2, p, _, _) volume of an LES frame."""
        dims = np.zeros(3, np.uint32)
        p = self.lib.RS_get_les_cpxx(self.H, frame, _ptr(dims))
        if not p:
            return None
        nx, ny, nz = (int(v) for v in dims)
        buf = (ctypes.c_char * (nx * ny * nz * 16)).from_address(p)
        return np.frombuffer(buf, dtype=np.float32).reshape(nz, ny, nx, 4).copy()

    def debris_flux_field(self, slot):
        """(inverse CDF, descriptor16) of the debris flux field of an LES slot, or None."""
        n = self.lib.RS_get_debris_flux_field(self.H, slot, None, 0, None)
        if n <= 0:
            return None
        icdf = np.zeros(n, np.float32)
        desc = np.zeros(16, np.float32)
        self.lib.RS_get_debris_flux_field(self.H, slot, _ptr(icdf), n, _ptr(desc))
        return icdf, desc

    def debris_table(self, t, which):
        """(ny, nx, 4) float32 texels of debris object t (0 cd, 1 cm, 2 rcs real, 3 rcs imag)."""
        dims = np.zeros(2, np.uint32)
        if self.lib.RS_get_debris_table(self.H, t, which, None, _ptr(dims)) != 0:
            return None
        nx, ny = int(dims[0]), int(dims[1])
        out = np.zeros((ny, nx, 4), np.float32)
        self.lib.RS_get_debris_table(self.H, t, which, _ptr(out), _ptr(dims))
        return out

    def dsd(self):
        cap = 64
        r = np.zeros(cap, np.float32)
        pdf = np.zeros(cap, np.float32)
        cdf = np.zeros(cap, np.float32)
        n = self.lib.RS_get_dsd(self.H, _ptr(r), _ptr(pdf), _ptr(cdf), cap)
        return r[:n].copy(), pdf[:n].copy(), cdf[:n].copy()

    def engine(self):
        return self.lib.RS_engine(self.H)

    def timer_start(self):
        e = self.engine()
        rc = self.lib.rsd_timer_start(e)
        if rc:
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def timer_stop(self):
        e = self.engine()
        ms = ctypes.c_double()
        rc = self.lib.rsd_timer_stop(e, ctypes.byref(ms))
        if rc:
            raise RuntimeError(self.lib.rsd_last_error(e).decode())
        return ms.value

    def profile(self, enable=True, stride=1):
        """Record HIP-event timings of every stride-th rain-kernel launch."""
        e = self.engine()
        if self.lib.rsd_profile(e, int(stride) if enable else 0):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def profile_read(self, kind=-1):
        """(total ms, launches) of the recorded rain-kernel launches of one kind."""
        e = self.engine()
        t, c = ctypes.c_double(), ctypes.c_int()
        if self.lib.rsd_profile_read(e, kind, ctypes.byref(t), ctypes.byref(c)):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())
        return t.value, c.value

    def set_launch(self, block, max_blocks):
        e = self.engine()
        if self.lib.rsd_set_launch(e, block, max_blocks):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def set_position_precision(self, hilo):
        """1: carry positions as binary32 hi/lo pairs (binary64 range for the phase)."""
        e = self.engine()
        if self.lib.rsd_set_position_precision(e, int(hilo)):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def pos_lo(self):
        """(N, 4) low parts of the hi/lo positions, reference order."""
        e = self.engine()
        n = int(self.info().num_scats)
        out = np.zeros((n, 4), np.float32)
        if self.lib.rsd_download_pos_lo(e, _ptr(out)):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())
        return out

    def set_debris_surrogate(self, debris_type, net):
        """Use a surrogate network (surrogate.Mlp, or None for the tables) for a debris type."""
        e = self.engine()
        if self.lib.rsd_set_debris_surrogate(e, debris_type, net.m if net is not None else None):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def set_moment_pulses(self, pulses):
        """Keep the last `pulses` pulses on the device for the moment estimator."""
        self.lib.RS_set_moment_pulses(self.H, pulses)
        self.check()

    def moments(self, pulses):
        """(R, 4) float32: Z dB, V m/s, ZDR dB, rho_hv over the last `pulses` pulses."""
        R = int(self.params().range_count)
        out = np.zeros((R, 4), np.float32)
        n = self.lib.RS_get_moments(self.H, pulses, _ptr(out))
        if n != R:
            raise RuntimeError("RS_get_moments failed")
        return out

    def set_fold(self, mode):
        """Gate-sum strategy: 0 per-lane, 1 wave binary64, 2 wave binary32."""
        e = self.engine()
        if self.lib.rsd_set_fold(e, mode):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def set_sort(self, interval):
        """Pulses between spatial reorders of the rain streams (0 = never)."""
        e = self.engine()
        if self.lib.rsd_set_sort(e, interval):
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def sort_interval(self):
        """Pulses between spatial reorders of the rain streams in effect."""
        return int(self.lib.rsd_get_sort(self.engine()))

    def sync(self):
        e = self.engine()
        rc = self.lib.rsd_sync(e)
        if rc:
            raise RuntimeError(self.lib.rsd_last_error(e).decode())

    def check(self):
        inf = self.info()
        if inf.last_error:
            raise RuntimeError("engine call failed (%d): %s" % (inf.last_error, self.lib.rsd_last_error(self.engine()).decode()))


class Batch:
    """Independent rain domains (Radar objects on one GPU) pulsed by one
    persistent launch per round (rsd_batch_*).  Per round: make_pulse() on every
    radar records its pulse, launch() runs them all, then each radar's
    advance_time() / advance_beam() as usual."""

    def __init__(self, radars):
        self.radars = list(radars)
        self.lib = self.radars[0].lib
        arr = (ctypes.c_void_p * len(self.radars))(*[r.engine() for r in self.radars])
        b = ctypes.c_void_p()
        rc = self.lib.rsd_batch_create(arr, len(self.radars), ctypes.byref(b))
        if rc:
            raise RuntimeError("rsd_batch_create failed (%d)" % rc)
        self.b = b
        self._handles = (ctypes.c_void_p * len(self.radars))(*[r.H for r in self.radars])

    def _chk(self, rc):
        if rc:
            raise RuntimeError("batch: %s" % self.lib.rsd_batch_last_error(self.b).decode())

    def launch(self):
        self._chk(self.lib.rsd_batch_launch(self.b))

    def sync(self):
        self._chk(self.lib.rsd_batch_sync(self.b))

    def set_launch(self, chunk, blocks_per_cu):
        self._chk(self.lib.rsd_batch_set_launch(self.b, chunk, blocks_per_cu))

    def launch_shape(self):
        c, g, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(self.lib.rsd_batch_get_launch(self.b, ctypes.byref(c), ctypes.byref(g), ctypes.byref(t)))
        return c.value, g.value, bool(t.value)

    def set_graph(self, on):
        self._chk(self.lib.rsd_batch_set_graph(self.b, 1 if on else 0))

    def timer_start(self):
        self._chk(self.lib.rsd_batch_timer_start(self.b))

    def timer_stop(self):
        ms = ctypes.c_double()
        self._chk(self.lib.rsd_batch_timer_stop(self.b, ctypes.byref(ms)))
        return ms.value

    def round(self):
        """One pulse of every domain (RS_batch_round): make_pulse, one launch,
        advance_time, advance_beam."""
        self._chk(self.lib.RS_batch_round(self._handles, len(self.radars), self.b))

    def free(self):
        if self.b:
            self.lib.rsd_batch_destroy(self.b)
            self.b = None
