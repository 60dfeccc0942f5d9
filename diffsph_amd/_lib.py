"""ctypes binding of the C ABI in include/diffsph.h (libdiffsph_amd.so).

The library is loaded from the package directory.  There is no fallback:
if the HIP library is missing every op raises, so a GPU run can never
silently fall back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiffsph_amd.so")

c_f32p = C.c_void_p  # device pointers are passed as void*


class Domain(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [("min", C.c_float * 2), ("max", C.c_float * 2), ("periodic", C.c_int32 * 2)]


class Grid(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [("ncell", C.c_int32 * 2), ("hcell", C.c_float), ("pad_", C.c_int32)]


class Pairs(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("nq", C.c_int64),
        ("nr", C.c_int64),
        ("offsets", C.c_void_p),
        ("nbr", C.c_void_p),
        ("qpos", C.c_void_p),
        ("rpos", C.c_void_p),
        ("qh", C.c_void_p),
        ("rh", C.c_void_p),
        ("h_uniform", C.c_float),
        ("mode", C.c_int32),
        ("domain", Domain),
        ("counts", C.c_void_p),
        ("slice_off", C.c_void_p),
        ("nbr16", C.c_void_p),
        ("nbr16_overflow", C.c_void_p),
        ("tile_slot", C.c_void_p),
        ("tile_runs", C.c_void_p),
        ("tile_n", C.c_void_p),
        ("tile_cap", C.c_int32),
        ("tile_pad_", C.c_int32),
    ]


class PairArrays(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("num_rows", C.c_int64),
        ("num_ref", C.c_int64),
        ("num_pairs", C.c_int64),
        ("offsets", C.c_void_p),
        ("j", C.c_void_p),
        ("kernels", C.c_void_p),
        ("gradients", C.c_void_p),
        ("distances", C.c_void_p),
        ("vectors", C.c_void_p),
        ("supports", C.c_void_p),
        ("laplacians", C.c_void_p),
    ]


class DfsphParams(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("dt", C.c_float),
        ("rho0", C.c_float),
        ("omega", C.c_float),
        ("error_threshold", C.c_float),
        ("min_iters", C.c_int32),
        ("max_iters", C.c_int32),
        ("clamp_pressure", C.c_int32),
        ("source_density", C.c_int32),
        ("check_every", C.c_int32),
        ("scheme", C.c_int32),
    ]


class SolveStats(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("final_error", C.c_double),
        ("initial_error", C.c_double),
    ]


class ViscParams(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("alpha", C.c_float),
        ("cs", C.c_float),
        ("rho0", C.c_float),
        ("eps", C.c_float),
        ("pi_switch", C.c_int32),
        ("scheme", C.c_int32),
    ]


class DeltaSPHParams(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("rho0", C.c_float),
        ("cs", C.c_float),
        ("delta", C.c_float),
        ("kernel_scale", C.c_float),
        ("visc_coeff", C.c_float),
        ("visc_eps", C.c_float),
        ("eos_kappa", C.c_float),
        ("gravity", C.c_float * 2),
        ("viscosity", C.c_int32),
        ("pi_switch", C.c_int32),
        ("eos", C.c_int32),
        ("diffusion", C.c_int32),
        ("pad_", C.c_int32),
    ]


class Boundary(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("pairs", C.POINTER(Pairs)),
        ("mass", C.c_void_p),
        ("rho", C.c_void_p),
        ("vel", C.c_void_p),
        ("visc_coeff", C.c_float),
        ("visc_eps", C.c_float),
        ("viscosity", C.c_int32),
        ("pi_switch", C.c_int32),
    ]


class RK4Stages(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("dxdt", C.c_void_p * 4),
        ("dudt", C.c_void_p * 4),
        ("drhodt", C.c_void_p * 4),
    ]


class BicgstabStats(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("iterations", C.c_int32),
        ("status", C.c_int32),
        ("residual_norm", C.c_double),
        ("rhs_norm", C.c_double),
    ]


class TimestepParams(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("support", C.c_float),
        ("kernel_scale", C.c_float),
        ("dt_visc", C.c_float),
        ("dt_acoustic", C.c_float),
        ("max_dt", C.c_float),
        ("min_dt", C.c_float),
        ("adaptive", C.c_int32),
        ("viscosity_constraint", C.c_int32),
        ("acceleration_constraint", C.c_int32),
        ("acoustic_constraint", C.c_int32),
    ]


class Batch(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("n", C.c_int64),
        ("ncell", C.c_int64),
        ("ndom", C.c_int32),
        ("width", C.c_int32),
        ("block", C.c_int32),
        ("support", C.c_float),
        ("hsearch", C.c_float),
        ("pad_", C.c_int32),
        ("workspace", C.c_void_p),
        ("workspace_bytes", C.c_size_t),
    ]


class ShiftParams(C.Structure):
    __slots__ = ()  # a misspelt field name raises instead of being ignored
    _fields_ = [
        ("dx", C.c_double),
        ("support", C.c_double),
        ("kernel_scale", C.c_float),
        ("R", C.c_float),
        ("n", C.c_float),
        ("cfl", C.c_float),
        ("cs", C.c_float),
        ("threshold", C.c_float),
        ("mach", C.c_float),
        ("pad_", C.c_int32),
    ]


V = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F32 = C.c_float
SZ = C.c_size_t
P = C.POINTER

# name -> (restype, argtypes)
SIGNATURES = {
    "dsph_status_string": (C.c_char_p, [C.c_int]),
    "dsph_abi_version": (C.c_int, []),
    "dsph_grid_dims": (C.c_int, [P(Domain), F32, P(Grid)]),
    "dsph_grid_build_workspace": (SZ, [I64, P(Grid)]),
    "dsph_grid_build": (C.c_int, [P(Domain), P(Grid), V, V, I64, V, V, V, V, V, V, SZ, V]),
    "dsph_neighbor_count": (C.c_int, [P(Domain), P(Grid), V, V, V, I64, V, V, I64, I32, F32, I32, V, V]),
    "dsph_neighbor_fill": (C.c_int, [P(Domain), P(Grid), V, V, V, I64, V, V, I64, I32, F32, I32, V, V, I64, V, I64,
                                     V]),
    "dsph_neighbor_fill16": (C.c_int, [P(Domain), P(Grid), V, V, V, I64, V, V, I64, I32, F32, I32, V, V, I64, V, V,
                                       I64, V, V]),
    "dsph_neighbor_fill_fixed": (C.c_int, [P(Domain), P(Grid), V, V, V, I64, V, V, I64, I32, F32, I32, I32, V, V,
                                           V, I64, V, V, V, V]),
    "dsph_neighbor_fill_fixed_map": (C.c_int, [P(Domain), P(Grid), V, V, V, I64, V, V, I64, I32, F32, I32, I32, V,
                                               V, V, I64, V, V, V]),
    "dsph_neighbor_fill_fixed_tiles": (C.c_int, [P(Domain), P(Grid), V, V, V, I64, V, V, I64, I32, F32, I32, I32,
                                                 V, V, V, I64, V, V, I32, V, I64, V, V, V, V]),
    "dsph_sell_workspace": (SZ, [I64]),
    "dsph_sell_layout": (C.c_int, [V, I64, V, P(I64), V, SZ, V]),
    "dsph_sell_to_csr": (C.c_int, [P(Pairs), V, V]),
    "dsph_wrap_positions": (C.c_int, [P(Domain), V, I64, V, V]),
    "dsph_neighbor_compact16": (C.c_int, [P(Pairs), I32, V, I64, V, V]),
    "dsph_neighbor_tiles": (C.c_int, [P(Pairs), P(Grid), V, I32, V, I64, V, V, V]),
    "dsph_scan_workspace": (SZ, [I64]),
    "dsph_scan_exclusive": (C.c_int, [V, I64, V, P(I64), V, SZ, V]),
    "dsph_gather_f32": (C.c_int, [V, V, I64, I32, V, V]),
    "dsph_gather_set": (C.c_int, [V, I64, I32, V, V, V, V]),
    "dsph_scatter_f32": (C.c_int, [V, V, I64, I32, V, V]),
    "dsph_pairs_export": (C.c_int, [P(Pairs), V, V, V, V, V, V, V, V, V, V, V, V, SZ, V]),
    "dsph_pairs_export_workspace": (SZ, [I64]),
    "dsph_density": (C.c_int, [P(Pairs), V, V, V, V, V]),
    "dsph_sph_operation": (C.c_int, [P(Pairs), I32, I32, V, V, V, V, V, V, I32, V, V]),
    "dsph_dfsph_prepare": (C.c_int, [P(Pairs), V, V, V, V, V, F32, F32, P(ViscParams), V, V, V, V, V, V]),
    "dsph_dfsph_workspace": (SZ, [I64]),
    "dsph_dfsph_solve": (C.c_int, [P(Pairs), P(DfsphParams), V, V, V, V, V, V, V, V, V, V, V, P(SolveStats), V, SZ,
                                   V]),
    "dsph_pressure_accel": (C.c_int, [P(Pairs), V, V, V, V, V, V]),
    "dsph_pressure_update": (C.c_int, [P(Pairs), P(DfsphParams), V, V, V, V, V, V, V, V, V]),
    "dsph_pack_pressure": (C.c_int, [P(Pairs), V, V, V]),
    "dsph_pressure_accel_rec": (C.c_int, [P(Pairs), V, V, V, V, V]),
    "dsph_pressure_update_rec": (C.c_int, [P(Pairs), P(DfsphParams), V, V, V, V, V, V, V, V, V]),
    "dsph_dfsph_iterate": (C.c_int, [P(Pairs), P(DfsphParams), V, V, V, V, V, V, V, V, I32, P(C.c_double), V, SZ,
                                     V]),
    "dsph_density_error": (C.c_int, [V, I64, F32, P(C.c_double), V, SZ, V]),
    "dsph_integrate_semi_implicit": (C.c_int, [V, V, V, V, V, V, V, I64, F32, V, V]),
    "dsph_ws_features": (C.c_int, [P(Pairs), V, V, V, V, V, V, F32, V, V, V, V, V, V]),
    "dsph_ws_features_ring": (C.c_int, [P(Pairs), V, V, V, V, V, V, F32, V, V, V, I32, V, V, V, V]),
    "dsph_density_w": (C.c_int, [P(Pairs), V, V, V, V, V, V]),
    "dsph_mlp_num_params": (C.c_int, []),
    "dsph_mlp_num_inputs": (C.c_int, []),
    "dsph_mlp_packed_bytes": (SZ, []),
    "dsph_mlp_pack": (C.c_int, [V, V, V]),
    "dsph_mlp_forward": (C.c_int, [V, I64, V, V, V, V]),
    "dsph_mlp_forward_affine": (C.c_int, [V, I64, V, V, V, F32, F32, I32, V, V]),
    "dsph_mlp_train_workspace": (SZ, [I64]),
    "dsph_mlp_loss_grad": (C.c_int, [V, V, I64, F32, V, V, V, V, V, SZ, V]),
    "dsph_adam_step": (C.c_int, [V, V, V, V, I64, F32, F32, F32, F32, F32, I32, V]),
    "dsph_remap_neighbors": (C.c_int, [V, I64, V, I64, V]),
    "dsph_deltasph_normalization": (C.c_int, [P(Pairs), P(Boundary), V, V, V, V, V, V]),
    "dsph_deltasph_rhs": (C.c_int, [P(Pairs), P(Boundary), P(DeltaSPHParams), V, V, V, V, V, V, V, V]),
    "dsph_volume": (C.c_int, [I64, V, V, V, V]),
    "dsph_deltasph_normalization_v": (C.c_int, [P(Pairs), P(Boundary), V, V, V, V, V, V, V]),
    "dsph_boundary_ghosts": (C.c_int, [V, V, V, I64, F32, V, V]),
    "dsph_mdbc": (C.c_int, [P(Pairs), V, V, V, F32, F32, I32, V, V, V]),
    "dsph_rk_update": (C.c_int, [I64, V, V, V, V, V, V, V, V, V, F32, V]),
    "dsph_rk4_combine": (C.c_int, [I64, P(RK4Stages), V, V, V, F32, V, V, V, V]),
    "dsph_rk_update_dev": (C.c_int, [I64, V, V, V, V, V, V, V, V, V, V, I32, V]),
    "dsph_rk4_combine_dev": (C.c_int, [I64, P(RK4Stages), V, V, V, V, V, V, V, V]),
    "dsph_timestep_dev": (C.c_int, [V, I64, P(TimestepParams), I32, V, V, V, SZ, V]),
    "dsph_shift_workspace": (SZ, [I64]),
    "dsph_deltaplus_shift": (C.c_int, [P(Pairs), P(Boundary), P(ShiftParams), V, V, V, V, V, V, SZ, V]),
    "dsph_max_norm_workspace": (SZ, []),
    "dsph_max_norm2": (C.c_int, [V, I64, P(C.c_float), V, SZ, V]),
    "dsph_sum_squares": (C.c_int, [V, I64, P(C.c_double), V, SZ, V]),
    "dsph_isph_laplacian": (C.c_int, [P(Pairs), V, V, F32, V, V, V, V]),
    "dsph_spmv": (C.c_int, [P(Pairs), V, V, V, V]),
    "dsph_bicgstab_workspace": (SZ, [I64]),
    "dsph_bicgstab": (C.c_int, [P(Pairs), V, V, V, V, F32, C.c_double, C.c_double, I32, I32, I32,
                                P(BicgstabStats), V, SZ, V]),
    "dsph_isph_source_workspace": (SZ, [I64]),
    "dsph_isph_source": (C.c_int, [P(Pairs), V, V, V, V, F32, F32, V, V, SZ, V]),
    "dsph_isph_jacobi_init": (C.c_int, [I64, V, F32, V, V, V, V]),
    "dsph_isph_close": (C.c_int, [P(Pairs), V, V, V, V, V, V, F32, F32, V, V, V, V, V]),
    "dsph_batch_workspace": (SZ, [P(Domain), P(I64), I32, F32, I32]),
    "dsph_batch_init": (C.c_int, [P(Domain), P(I64), I32, F32, F32, I32, I32, V, SZ, P(Batch)]),
    "dsph_batch_step": (C.c_int, [P(Batch), P(DeltaSPHParams), P(TimestepParams), V, V, V, V, V, V, V, I32, V]),
    "dsph_batch_graph": (C.c_int, [P(Batch), P(DeltaSPHParams), P(TimestepParams), V, V, V, V, V, V, V, I32,
                                   P(C.c_void_p)]),
    "dsph_graph_launch": (C.c_int, [V, V]),
    "dsph_graph_destroy": (C.c_int, [V]),
    "dsph_batch_replay_pass": (C.c_int, [P(Batch), P(DeltaSPHParams), I32, V]),
    "dsph_batch_probe": (C.c_int, [P(Batch), V, P(I32), P(I64), V]),
    "dsph_batch_status": (C.c_int, [P(Batch), P(I32), I32, V]),
    "dsph_deltasph_vjp_workspace": (SZ, [I64]),
    "dsph_deltasph_rhs_vjp": (C.c_int, [P(Pairs), P(DeltaSPHParams), V, V, V, V, V, V, V, V, V, V, V, SZ, V]),
    "dsph_op_sph_density": (C.c_int, [V, I64, V, V, I64, I32, V, V, V, V]),
    "dsph_op_sph_interpolation": (C.c_int, [V, V, V, I32, I64, V, V, I64, I32, V, V, V, V]),
    "dsph_op_scatter_add": (C.c_int, [V, I32, I64, I32, V, V, V, V]),
    "dsph_op_viscosity": (C.c_int, [V, V, V, V, V, V, I64, V, V, V, V, V, I64, I32, V, V, F32, F32, F32, F32, I32,
                                    V, V]),
    "dsph_op_mdbc": (C.c_int, [V, V, I64, V, V, I32, V, V, V, V, V, V, I64, V, V, F32, V, V, V]),
    "dsph_op_neighborhood_count": (C.c_int, [V, V, I32, V, V, I64, V, I64, V, V, V, V, V, V, V]),
    "dsph_op_neighborhood_fill": (C.c_int, [V, V, I32, V, V, I64, V, I64, V, V, V, V, V, V, I64, V, V, V, V, V,
                                            V]),
    "dsph_pairs_filter_count": (C.c_int, [P(Domain), I32, V, V, I32, V, V, I64, V, I64, V, V, V, V]),
    "dsph_pairs_filter_fill": (C.c_int, [P(Domain), I32, V, V, I32, V, V, I64, V, I64, V, V, V, I64, V, V, V, V,
                                         V, V, V, V]),
    "dsph_verlet_check_workspace": (SZ, []),
    "dsph_verlet_check": (C.c_int, [V, V, I64, V, I64, P(C.c_float), V, SZ, V]),
    "dsph_pp_offsets": (C.c_int, [V, I64, I32, V, V, V]),
    "dsph_pp_operation": (C.c_int, [P(PairArrays), I32, I32, I32, V, V, V, V, V, V, V, I32, V, V]),
    "dsph_pp_dfsph_alpha": (C.c_int, [P(PairArrays), V, V, V, V, F32, F32, V, V]),
    "dsph_pp_dfsph_update_workspace": (SZ, [I64]),
    "dsph_pp_dfsph_update": (C.c_int, [P(PairArrays), V, V, V, V, V, V, V, F32, F32, I32, F32, F32, V, V, V, V, SZ,
                                       V]),
    "dsph_ips_assemble": (C.c_int, [P(Pairs), V, V, V, I64, V]),
    "dsph_ips_spmv": (C.c_int, [P(Pairs), V, I64, V, V, V]),
    "dsph_vec_lincomb": (C.c_int, [I64, V, F32, V, F32, V, V, V]),
    "dsph_vec_reduce_workspace": (SZ, [I64]),
    "dsph_vec_dot": (C.c_int, [I64, V, V, V, V, SZ, V]),
    "dsph_vec_max_norm2": (C.c_int, [I64, V, V, V, SZ, V]),
    "dsph_noalloc_scratch_bytes": (SZ, []),
    "dsph_dfsph_solve_noalloc_max_particles": (I64, []),
    "dsph_dfsph_solve_noalloc": (C.c_int, [P(Pairs), P(DfsphParams), V, V, V, V, V, V, V, V, V, V, V, P(SolveStats),
                                           V]),
    "dsph_hash_count_neighbors": (C.c_int, [V, V, I32, I32, V, V, I32, V, I32, V, V, I64, F32, V, V, V, I32, V,
                                            V]),
    "dsph_hash_build_neighbors": (C.c_int, [V, I64, V, V, I32, I32, V, V, I32, V, I32, V, V, I64, F32, V, V, V,
                                            I32, V, V, V]),
}

_lib = None


class DsphError(RuntimeError):
    pass


def lib():
    """Load (once) and return the HIP library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DsphError(
                f"HIP library {LIB_PATH} is not built; run `python -m diffsph_amd.build` "
                "(the diffsph_amd ops have no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().dsph_status_string(status).decode()
        raise DsphError(f"{what}: {msg} (status {status})")


def ptr(t) -> int | None:
    """Raw device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def make_domain(min_extent, max_extent, periodic) -> Domain:
    d = Domain()
    mn = [float(v) for v in (min_extent if hasattr(min_extent, "__len__") else [min_extent] * 2)]
    mx = [float(v) for v in (max_extent if hasattr(max_extent, "__len__") else [max_extent] * 2)]
    pr = periodic if hasattr(periodic, "__len__") else [periodic] * 2
    for k in range(2):
        d.min[k] = mn[k]
        d.max[k] = mx[k]
        d.periodic[k] = 1 if bool(pr[k]) else 0
    return d
