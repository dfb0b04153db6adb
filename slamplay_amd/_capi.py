"""ctypes declarations of the C ABI in include/mba/ba_c_api.h (libslamplay_ba.so).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libslamplay_ba.so")

i32 = ctypes.c_int32
i64 = ctypes.c_int64
f64 = ctypes.c_double
i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p

BA_OK = 0
STATUS_NAMES = {0: "BA_OK", 1: "BA_ERR_INVALID_ARGUMENT", 2: "BA_ERR_OUT_OF_MEMORY", 3: "BA_ERR_NON_FINITE",
                4: "BA_ERR_HIP", 5: "BA_ERR_NO_DEVICE", 6: "BA_ERR_CAPACITY", 7: "BA_ERR_IO", 8: "BA_ERR_NOT_FOUND",
                9: "BA_ERR_TIMEOUT"}
PRECISION_FP64, PRECISION_MIXED = 0, 1
SOLVER_AUTO, SOLVER_DENSE_SCHUR, SOLVER_ITERATIVE_SCHUR = 0, 1, 2
LOSS_NONE, LOSS_HUBER = 0, 1
CONVERGENCE, NO_CONVERGENCE, FAILURE = 0, 1, 2


class ProblemView(ctypes.Structure):
    _fields_ = [("num_cameras", i32), ("num_points", i32), ("num_observations", i32), ("camera_block_size", i32),
                ("camera_index", i32p), ("point_index", i32p), ("observations", f64p), ("cameras", f64p),
                ("points", f64p)]


class Options(ctypes.Structure):
    _fields_ = [("max_num_iterations", i32), ("max_num_consecutive_invalid_steps", i32), ("jacobi_scaling", i32),
                ("loss", i32), ("linear_solver", i32), ("pcg_max_iterations", i32), ("precision", i32),
                ("dense_schur_max_cameras", i32), ("device", i32), ("verbose", i32), ("refine_steps", i32),
                ("launch_mode", i32), ("loss_scale", f64), ("function_tolerance", f64), ("gradient_tolerance", f64),
                ("parameter_tolerance", f64), ("initial_trust_region_radius", f64), ("max_trust_region_radius", f64),
                ("min_trust_region_radius", f64), ("min_relative_decrease", f64), ("min_lm_diagonal", f64),
                ("max_lm_diagonal", f64), ("pcg_tolerance", f64), ("mode", i32), ("g2o_max_trials", i32),
                ("g2o_tau", f64)]


class Iteration(ctypes.Structure):
    _fields_ = [("iteration", i32), ("step_is_successful", i32), ("step_is_valid", i32),
                ("linear_solver_iterations", i32), ("cost", f64), ("cost_change", f64), ("gradient_max_norm", f64),
                ("step_norm", f64), ("relative_decrease", f64), ("trust_region_radius", f64),
                ("model_cost_change", f64), ("iteration_time_in_seconds", f64), ("cumulative_time_in_seconds", f64)]


class Summary(ctypes.Structure):
    _fields_ = [("initial_cost", f64), ("final_cost", f64), ("num_iterations", i32), ("num_successful_steps", i32),
                ("num_unsuccessful_steps", i32), ("termination_type", i32), ("linear_solver_used", i32),
                ("total_linear_solver_iterations", i32), ("total_time_ms", f64), ("trace", ctypes.POINTER(Iteration)),
                ("trace_capacity", i32), ("termination_reason", i32), ("termination_value", f64)]


class BatchSchedule(ctypes.Structure):
    _fields_ = [("schur_threads", i32), ("solve_threads", i32), ("blocks_per_wave", i32), ("lds_cameras", i32),
                ("candidates", i32), ("iteration_ms", f64), ("default_iteration_ms", f64), ("lanes", i32)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", i64), ("total_ms", f64), ("bytes_per_launch", f64),
                ("flops_per_launch", f64)]


class SynthParams(ctypes.Structure):
    _fields_ = [("kind", i32), ("num_cameras", i32), ("num_points", i32), ("min_track", i32), ("max_track", i32),
                ("target_observations", i32), ("zipf_s", f64), ("pixel_noise", f64), ("rotation_sigma", f64),
                ("translation_sigma", f64), ("point_sigma", f64), ("focal_mean", f64), ("focal_sigma", f64),
                ("distortion_sigma", f64), ("seed", ctypes.c_uint64)]


class StereoView(ctypes.Structure):
    _fields_ = [("num_poses", i32), ("num_points", i32), ("num_observations", i32), ("pose_index", i32p),
                ("point_index", i32p), ("side", i32p), ("observations", f64p), ("poses", f64p), ("points", f64p),
                ("K", f64 * 4), ("extrinsics", (f64 * 6) * 2)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("arena_bytes", i64), ("max_windows", i32), ("iterations_per_step", i32), ("options", Options)]


class ServerResult(ctypes.Structure):
    _fields_ = [("tag", i64), ("initial_cost", f64), ("final_cost", f64), ("iterations", i32),
                ("termination_type", i32), ("submit_time_s", f64), ("admit_time_s", f64), ("finish_time_s", f64),
                ("num_cameras", i32), ("num_points", i32), ("cameras", f64p), ("points", f64p),
                ("evictions", i32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "ba_abi_version": (i32, []),
    "ba_last_error_string": (ctypes.c_char_p, []),
    "ba_default_options": (None, [ctypes.POINTER(Options)]),
    "ba_device_count": (ctypes.c_int, [i32p]),
    "ba_eval_residual_jacobian": (ctypes.c_int, [ctypes.POINTER(ProblemView), i32, f64p, f64p, f64p]),
    "ba_eval_cost": (ctypes.c_int, [ctypes.POINTER(ProblemView), i32, f64, f64p]),
    # device pointers throughout (the view's arrays, the one-double cost) and a hipStream_t
    "ba_eval_cost_noalloc": (ctypes.c_int, [ctypes.POINTER(ProblemView), i32, f64, vp, vp]),
    "ba_solve": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.POINTER(Options), ctypes.POINTER(Summary)]),
    "ba_summary_full_report": (ctypes.c_int, [ctypes.POINTER(Summary), i32, i32, i32, ctypes.POINTER(Options),
                                               ctypes.c_char_p, i64, ctypes.POINTER(i64)]),
    "ba_solver_create": (ctypes.c_int, [ctypes.POINTER(ProblemView), ctypes.POINTER(Options), ctypes.POINTER(vp)]),
    "ba_solver_run": (ctypes.c_int, [vp, ctypes.POINTER(Summary)]),
    "ba_solver_reset": (ctypes.c_int, [vp]),
    "ba_solver_download": (ctypes.c_int, [vp, f64p, f64p]),
    "ba_solver_normalize": (ctypes.c_int, [vp]),
    "ba_solver_perturb": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "ba_solver_destroy": (ctypes.c_int, [vp]),
    "ba_solver_set_profiling": (ctypes.c_int, [vp, i32]),
    "ba_solver_kernel_times": (ctypes.c_int, [vp, ctypes.POINTER(KernelTime), i32, i32p]),
    "ba_solver_info": (ctypes.c_int, [vp, ctypes.POINTER(i64), i32p]),
    "ba_batch_create": (ctypes.c_int, [ctypes.POINTER(ProblemView), i32, ctypes.POINTER(Options),
                                        ctypes.POINTER(vp)]),
    "ba_batch_run": (ctypes.c_int, [vp, ctypes.POINTER(Summary)]),
    "ba_batch_reset": (ctypes.c_int, [vp]),
    "ba_batch_download": (ctypes.c_int, [vp, i32, f64p, f64p]),
    "ba_batch_destroy": (ctypes.c_int, [vp]),
    "ba_batch_solve": (ctypes.c_int, [ctypes.POINTER(ProblemView), i32, ctypes.POINTER(Options),
                                       ctypes.POINTER(Summary)]),
    "ba_batch_set_block": (ctypes.c_int, [vp, i32]),
    "ba_batch_get_schedule": (ctypes.c_int, [vp, ctypes.POINTER(BatchSchedule)]),
    "ba_batch_set_schedule": (ctypes.c_int, [vp, ctypes.POINTER(BatchSchedule)]),
    "ba_batch_info": (ctypes.c_int, [vp, i32p, ctypes.POINTER(i64)]),
    "ba_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
    "ba_server_submit": (ctypes.c_int, [vp, ctypes.POINTER(ProblemView), i64]),
    "ba_server_step": (ctypes.c_int, [vp, i32p, i32p]),
    "ba_server_poll": (ctypes.c_int, [vp, ctypes.POINTER(ServerResult), i32, i32p]),
    "ba_server_cancel": (ctypes.c_int, [vp, i64]),
    "ba_server_evict": (ctypes.c_int, [vp, i64]),
    "ba_server_clock": (f64, [vp]),
    "ba_server_destroy": (ctypes.c_int, [vp]),
    "ba_server_simulate": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(ProblemView), i32, f64, i32,
                                           ctypes.c_uint64, ctypes.POINTER(ServerResult)]),
    "ba_stereo_default_options": (None, [ctypes.POINTER(Options)]),
    "ba_stereo_batch_solve": (ctypes.c_int, [ctypes.POINTER(StereoView), i32, ctypes.POINTER(Options),
                                              ctypes.POINTER(Summary), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                              f64p]),
    "ba_stereo_solve": (ctypes.c_int, [ctypes.POINTER(StereoView), ctypes.POINTER(Options), ctypes.POINTER(Summary),
                                        ctypes.POINTER(ctypes.c_uint8), f64p]),
    "ba_bal_load": (ctypes.c_int, [ctypes.c_char_p, i32, ctypes.POINTER(vp)]),
    "ba_bal_from_arrays": (ctypes.c_int, [i32, i32, i32, i32p, i32p, f64p, f64p, f64p, ctypes.POINTER(vp)]),
    "ba_bal_view": (ctypes.c_int, [vp, ctypes.POINTER(ProblemView)]),
    "ba_bal_normalize": (ctypes.c_int, [vp]),
    "ba_bal_perturb": (ctypes.c_int, [vp, f64, f64, f64]),
    "ba_bal_write": (ctypes.c_int, [vp, ctypes.c_char_p, i32]),
    "ba_bal_write_ply": (ctypes.c_int, [vp, ctypes.c_char_p]),
    "ba_bal_destroy": (ctypes.c_int, [vp]),
    "ba_synth_default": (ctypes.c_int, [i32, ctypes.POINTER(SynthParams)]),
    "ba_synth_generate": (ctypes.c_int, [ctypes.POINTER(SynthParams), ctypes.POINTER(vp), f64p, i64]),
}

_lib = None


def lib():
    """Load libslamplay_ba.so (built in-tree by __graft_entry__.build() / make)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libslamplay_ba.so is not built ({LIB_PATH}); run `make -C slamplay_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class BAError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status):
    if status != BA_OK:
        msg = lib().ba_last_error_string()
        raise BAError(status, msg.decode() if msg else "")
    return status
