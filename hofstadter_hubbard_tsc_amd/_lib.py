"""ctypes binding of libhhdmrg.so (the C ABI in include/hhdmrg.h).

The library is built in-tree (``python -m hofstadter_hubbard_tsc_amd.build``);
loading fails loudly when it is missing -- there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HHDMRG_LIB", os.path.join(HERE, "libhhdmrg.so"))

i32, i64, f64, u64 = C.c_int32, C.c_int64, C.c_double, C.c_uint64
vp = C.c_void_p
P = C.POINTER


class ModelParams(C.Structure):
    _fields_ = [("kind", i32), ("Lx", i32), ("Ly", i32), ("lattice", i32), ("periodic_y", i32),
                ("n_elec", i32), ("t", f64), ("U", f64), ("J", f64), ("flux", f64), ("gauge", i32),
                ("reserved", i32), ("theta", f64)]


class Schedule(C.Structure):
    _fields_ = [("kept_max", i32), ("kept_max_warm", i32), ("n_sweeps", i32), ("max_iter", i32),
                ("tol", f64), ("warm_tol_factor", f64), ("first_tol_factor", f64), ("seed", u64),
                ("precision", i32), ("solver", i32)]


class StepRecord(C.Structure):
    _fields_ = [("phase", i32), ("sys_len", i32), ("env_len", i32), ("iters", i32), ("num_state", i32),
                ("converged", i32), ("dim", i64), ("energy", f64), ("trunc_err", f64), ("entropy", f64),
                ("seconds", f64), ("matvec_seconds", f64), ("matvec_flops", f64),
                ("bind_seconds", f64), ("wave_seconds", f64), ("lanczos_seconds", f64), ("trunc_seconds", f64),
                ("grow_seconds", f64), ("ritz_change", f64), ("residual", f64), ("hpsi_fp64", i32),
                ("hpsi_fp32", i32), ("solver", i32), ("fallback", i32)]


class Stats(C.Structure):
    _fields_ = [("matvecs", i64), ("matvec_ms", f64), ("stage1_ms", f64), ("stage2_ms", f64),
                ("flops_alg", f64), ("flops_exec", f64), ("bytes_alg", f64), ("stage1_flops_alg", f64),
                ("stage2_flops_alg", f64), ("reduce_ms", f64), ("stage1_flops_exec", f64),
                ("stage2_flops_exec", f64), ("matvecs_fp32", i64), ("stage2_ms_fp32", f64),
                ("stage2_flops_alg_fp32", f64)]


# numpy views of the engine's task structs (hh_core.hpp: Term, Tile)
TERM_DTYPE = np.dtype([("cr", "<f8"), ("ci", "<f8"), ("a_off", "<i8"), ("b_off", "<i8"), ("a_ld", "<i4"),
                       ("b_ld", "<i4"), ("K", "<i4"), ("a_buf", "u1"), ("b_buf", "u1"), ("a_op", "u1"),
                       ("b_op", "u1")], align=True)
TILE_DTYPE = np.dtype([("c_off", "<i8"), ("c_ld", "<i4"), ("M", "<i4"), ("N", "<i4"), ("t0", "<i4"),
                       ("t1", "<i4"), ("c_buf", "<i4"), ("accumulate", "<i4")], align=True)

# (name, restype, argtypes): every entry point of include/hhdmrg.h
SIGNATURES = [
    ("hh_ctx_create", i32, [i32, P(vp)]),
    ("hh_ctx_destroy", i32, [vp]),
    ("hh_last_error", C.c_char_p, [vp]),
    ("hh_set_model", i32, [vp, P(ModelParams)]),
    ("hh_boundary_sites", i32, [vp, i32, P(i32), P(i32)]),
    ("hh_dev_alloc", i32, [vp, i64, P(vp)]),
    ("hh_dev_free", i32, [vp, vp]),
    ("hh_memcpy_h2d", i32, [vp, vp, vp, i64]),
    ("hh_memcpy_d2h", i32, [vp, vp, vp, i64]),
    ("hh_block_site", i32, [vp, i32, P(vp)]),
    ("hh_block_create", i32, [vp, i32, i32, i32, P(i32), P(i32), P(i32), i32, P(i32), P(vp)]),
    ("hh_block_destroy", i32, [vp]),
    ("hh_block_info", i32, [vp, P(i64), P(i32), P(i32), P(i32)]),
    ("hh_block_sectors", i32, [vp, P(i32), P(i32), P(i32), P(i64)]),
    ("hh_block_sites", i32, [vp, P(i32)]),
    ("hh_block_layout", i32, [vp, i32, P(i32), P(i32), P(i32), P(i64)]),
    ("hh_block_op_offset", i32, [vp, i32, i32, P(i64)]),
    ("hh_block_upload", i32, [vp, vp, vp]),
    ("hh_block_download", i32, [vp, vp, vp]),
    ("hh_superblock_bind", i32, [vp, vp, vp, i32, P(vp)]),
    ("hh_superblock_destroy", i32, [vp]),
    ("hh_superblock_info", i32, [vp, P(i64), P(i32), P(f64), P(f64)]),
    ("hh_superblock_pairs", i32, [vp, P(i32), P(i32), P(i32), P(i32), P(i64)]),
    ("hh_superblock_enlarged", i32, [vp, i32, P(i32), P(i32), P(i32), P(i32)]),
    ("hh_model_wave", i32, [vp, vp, vp, vp]),
    ("hh_batch_create", i32, [vp, i32, P(vp), P(vp)]),
    ("hh_batch_destroy", i32, [vp]),
    ("hh_batch_info", i32, [vp, P(i64), P(f64)]),
    ("hh_batch_model_wave", i32, [vp, vp, P(vp), P(vp)]),
    ("hh_batch_lanczos", i32, [vp, vp, P(vp), f64, i32, P(f64), P(i32), P(i32)]),
    ("hh_batch_set_graph", i32, [vp, i32]),
    ("hh_set_autotune", i32, [vp, i32]),
    ("hh_autotune_stats", i32, [vp, P(i32), P(i32), P(i32), P(f64)]),
    ("hh_set_launch_config", i32, [vp, i32, i32, i32, i32, i32]),
    ("hh_lanczos", i32, [vp, vp, vp, f64, i32, P(f64), P(i32)]),
    ("hh_lanczos_levels", i32, [vp, vp, P(vp), i32, f64, i32, P(f64), P(i32)]),
    ("hh_density_matrix_levels", i32, [vp, vp, i32, vp, vp, vp]),
    ("hh_truncate_levels", i32, [vp, vp, i32, vp, vp, i32, i32, f64, P(vp), P(i32), P(f64), P(f64)]),
    ("hh_lanczos_status", i32, [vp, P(i32), P(f64), P(f64)]),
    ("hh_truncate", i32, [vp, vp, i32, vp, i32, i32, f64, P(vp), P(i32), P(f64), P(f64)]),
    ("hh_trunc_identity", i32, [vp, vp, i32, P(vp)]),
    ("hh_trunc_destroy", i32, [vp]),
    ("hh_trunc_kept", i32, [vp, P(i32), P(i32), P(i32)]),
    ("hh_grow_block", i32, [vp, vp, i32, vp, P(vp)]),
    ("hh_renormalize", i32, [vp, vp, i32, P(i32), P(i32), P(i64), vp, P(i32), P(i64), P(i32), P(i32), P(i64), vp]),
    ("hh_block_conjugate", i32, [vp, vp]),
    ("hh_wave_trans", i32, [vp, i32, vp, vp, vp, vp, vp, vp]),
    ("hh_dmrg_run", i32, [vp, P(Schedule), P(StepRecord), i32, P(i32)]),
    ("hh_driver_create", i32, [vp, P(Schedule), P(vp)]),
    ("hh_driver_destroy", i32, [vp]),
    ("hh_driver_warmup", i32, [vp, P(StepRecord), i32, P(i32)]),
    ("hh_driver_sweep", i32, [vp, i32, i32, i32, f64, f64, i32, i32, i32, P(StepRecord), i32, P(i32)]),
    ("hh_driver_reset_transform", i32, [vp]),
    ("hh_driver_wave", i32, [vp, P(i32), P(i32), P(i64), vp]),
    ("hh_driver_set_solver", i32, [vp, i32]),
    ("hh_driver_fork", i32, [vp, P(vp)]),
    ("hh_eigh_sectors", i32, [vp, i32, P(i32), P(i32), vp, vp, i32]),
    ("hh_set_eigensolver", i32, [vp, i32]),
    ("hh_eigensolver_status", i32, [vp, P(i32), P(i64), P(i32), P(f64)]),
    ("hh_driver_fork_window", i32, [vp, i32, i32, i32, i32, P(vp)]),
    ("hh_driver_capture", i32, [vp, i32]),
    ("hh_driver_train_precond", i32, [vp, i32, f64, P(f64), P(i32)]),
    ("hh_driver_set_precond", i32, [vp, i32]),
    ("hh_driver_copy_precond", i32, [vp, vp]),
    ("hh_driver_set_min_dim", i32, [vp, i64]),
    ("hh_driver_set_subspace", i32, [vp, i32]),
    ("hh_driver_set_levels", i32, [vp, i32]),
    ("hh_driver_levels", i32, [vp, P(i32)]),
    ("hh_driver_set_level_start", i32, [vp, i32]),
    ("hh_driver_level_energies", i32, [vp, i32, P(f64), P(i32)]),
    ("hh_driver_set_fallback", i32, [vp, f64]),
    ("hh_precond_eval", i32, [vp, i32, P(i32), P(i32), P(i32), P(i32), i32, P(f64), P(f64), P(f64), P(f64), f64,
                              P(C.c_float), P(C.c_float), P(C.c_float), P(f64), P(f64), P(C.c_float)]),
    ("hh_precond_size", i32, []),
    ("hh_driver_block_count", i32, [vp, i32, P(i32)]),
    ("hh_driver_block", i32, [vp, i32, i32, P(vp)]),
    ("hh_set_profiling", i32, [vp, i32]),
    ("hh_set_precision", i32, [vp, i32]),
    ("hh_driver_density", i32, [vp, i32, vp, P(f64)]),
    ("hh_driver_density_cross", i32, [vp, i32, vp, P(f64)]),
    ("hh_driver_density_row", i32, [vp, i32, i32, vp, P(f64)]),
    ("hh_driver_save", i32, [vp, C.c_char_p]),
    ("hh_driver_load", i32, [vp, C.c_char_p]),
    ("hh_driver_spin_cross", i32, [vp, i32, vp, P(f64)]),
    ("hh_driver_spin_row", i32, [vp, i32, i32, vp, P(f64)]),
    ("hh_driver_hopping_cross", i32, [vp, i32, vp, P(f64)]),
    ("hh_driver_hopping_row", i32, [vp, i32, i32, vp, P(f64)]),
    ("hh_driver_pair_cross", i32, [vp, i32, vp, i32, P(i32), P(f64)]),
    ("hh_driver_pair_block", i32, [vp, i32, vp, i32, P(i32), P(f64)]),
    ("hh_driver_density_all", i32, [vp, i32, vp, P(f64)]),
    ("hh_density_matrix_layout", i32, [vp, i32, P(i32), P(i64), P(i32), P(i32), P(i64)]),
    ("hh_density_matrix", i32, [vp, vp, i32, vp, vp]),
    ("hh_wave_to_disk", i32, [vp, vp, vp, C.c_char_p]),
    ("hh_wave_from_disk", i32, [vp, vp, C.c_char_p, vp]),
    ("hh_trunc_to_disk", i32, [vp, vp, i32, C.c_char_p]),
    ("hh_trunc_from_disk", i32, [vp, vp, i32, C.c_char_p, P(i32), P(vp)]),
    ("hh_basis_to_disk", i32, [vp, vp, C.c_char_p]),
    ("hh_basis_from_disk", i32, [vp, vp, C.c_char_p, P(i32)]),
    ("hh_driver_spin_all", i32, [vp, i32, vp, P(f64)]),
    ("hh_driver_hopping_all", i32, [vp, i32, vp, P(f64)]),
    ("hh_driver_trunc", i32, [vp, i32, i32, P(vp)]),
    ("hh_trunc_layout", i32, [vp, P(i64), P(i64), P(i32)]),
    ("hh_trunc_download", i32, [vp, vp, vp]),
    ("hh_get_stats", i32, [vp, P(Stats)]),
    ("hh_reset_stats", i32, [vp]),
    ("hh_superblock_tasks", i32, [vp, i32, P(vp), P(i32), P(vp), P(i32), P(i64)]),
    ("hh_plan_matvec", i32, [P(ModelParams), i32, i32, P(i32), P(i32), P(i32), i32, P(i32), i32, i32, P(i32),
                             P(i32), P(i32), i32, P(i32), i32, P(vp)]),
    ("hh_plan_grow", i32, [P(ModelParams), i32, i32, i32, P(i32), P(i32), P(i32), i32, P(i32), i32, i32, i32,
                           P(i32), P(i32), P(i64), P(i32), P(vp)]),
    ("hh_plan_destroy", i32, [vp]),
    ("hh_plan_tasks", i32, [vp, i32, P(vp), P(i32), P(vp), P(i32), P(i64)]),
    ("hh_plan_block", i32, [vp, i32, P(i64), P(i32), P(i32), P(i32), P(i32), P(i64), P(i32), P(i32)]),
    ("hh_plan_layout", i32, [vp, i32, i32, P(i32), P(i32), P(i32), P(i64)]),
    ("hh_plan_op_offset", i32, [vp, i32, i32, i32, P(i64)]),
    ("hh_plan_enlarged", i32, [vp, i32, P(i32), P(i32), P(i32), P(i32), P(i32), P(i32)]),
    ("hh_plan_super", i32, [vp, P(i64), P(i32), P(i32), P(i32), P(i32), P(i32), P(i64)]),
    ("hh_run_tasks", i32, [vp, vp, i32, vp, i32, P(vp), i32]),
    ("hh_run_tasks_noalloc", i32, [vp, vp, i32, vp, i32, P(vp), i32]),
    ("hh_start_vector", i32, [vp, vp, i64, u64]),
]

_lib = None


def load():
    """Load libhhdmrg.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libhhdmrg.so is not built: run `python -m hofstadter_hubbard_tsc_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ip(a):
    """ctypes pointer to a contiguous numpy array."""
    if a is None:
        return None
    t = {np.dtype("int32"): i32, np.dtype("int64"): i64, np.dtype("float64"): f64}[a.dtype]
    return a.ctypes.data_as(P(t))


def model_params(m):
    """ModelParams from an object with kind/Lx/Ly/lattice/periodic_y/n_elec/t/U/J/flux (and optionally
    gauge "landau" | "reference" and theta)."""
    kind = {"hubbard": 0, "tj": 1}.get(m.kind, m.kind)
    lattice = {"square": 0, "triangular": 1}.get(m.lattice, m.lattice)
    g = getattr(m, "gauge", "landau")
    gauge = {"landau": 0, "reference": 1}.get(g, g)
    return ModelParams(kind, m.Lx, m.Ly, lattice, int(m.periodic_y), m.n_elec, m.t, m.U, m.J, m.flux, gauge, 0,
                       float(getattr(m, "theta", 0.0)))
