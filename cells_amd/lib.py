"""ctypes binding of libcells_hip.so (include/cells_hip.h).

The library is built in-tree by cells_amd.build (hipcc for gfx950). Loading fails loudly when it
is missing: the HIP step has no CPU fallback.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcells_hip.so")

CELLS_OK = 0
CELLS_EINVAL = 1
CELLS_EHIP = 2
CELLS_ENOMEM = 3
CELLS_ENODEV = 4
METHOD_EULER = 0
METHOD_ROS2 = 1
METHOD_N2 = 2  # the reference's ARS(2,2,2) with implicit slope conductances (N2)
METHOD_ARS222 = 3  # the reference's ARS(2,2,2) (its implicit part empty on a held bath)
METHOD_RK4 = 4  # the reference's RK4 of an empty tableau (worlds without water)
FACTOR_PAIRWISE = 0x100  # or-ed into the method: the reference-order factor

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)

EXPORTS = (
    "cells_last_error", "cells_minimum_degree", "cells_bind", "cells_symbolic_sizes",
    "cells_symbolic_export", "cells_symbolic_free", "cells_ensemble_create",
    "cells_ensemble_destroy", "cells_state_size", "cells_step", "cells_rates",
    "cells_factor_size", "cells_factor", "cells_solve", "cells_world_positions",
    "cells_kernel_times", "cells_device_checks", "cells_step_vjp", "cells_vmem_loss",
    "cells_set_parameters", "cells_plan_stats", "cells_tape_reserve", "cells_step_record",
    "cells_step_vjp_recorded", "cells_step_checks", "cells_step_check_bits", "cells_stationary",
    "cells_stationary_residual", "cells_stationary_jvp", "cells_field_create", "cells_field_destroy",
    "cells_field_state_size", "cells_field_step", "cells_field_refresh", "cells_field_plan_stats",
    "cells_field_kernel_times", "cells_field_set_start", "cells_field_potentials", "cells_field_energy_rates",
    "cells_world_factor_positions", "cells_relaxed_padding", "cells_set_plan_options",
    "cells_get_plan_options", "cells_step_bounds", "cells_set_stage_tolerance", "cells_step_vjp_noalloc", "cells_energy_rates",
    "cells_product_plan_size", "cells_product_plan", "cells_product", "cells_capacitor_topology",
)

# refusal reasons in a world's failure bits (cells_step_check_bits)
REFUSE_RESIDUAL = 1 << 16
REFUSE_EXPLICIT_RATE = 1 << 17

# the device row checks' failure bits (cells_device_checks), the reference's check names
CHECKS = {
    "domain": 1, "local": 2, "total": 4, "global_charge": 8, "dissipation": 16, "water": 32,
    "outside": 64, "charge": 128, "electrode": 256, "energy": 512, "reservoir_work": 1024,
    "electrode_work": 2048, "metabolic_work": 4096, "bath_water_work": 8192,
}
# the readouts' energy rates, in their column order (cells_energy_rates)
ENERGY_RATES = ("chemical", "capacitor", "dissipation", "reservoir_power", "electrode_power", "metabolic",
                "energy_balance", "electrode_current")
# the field solve's energy terms, in their column order (cells_field_energy_rates)
FIELD_ENERGY_RATES = ("chemical", "capacitor", "dissipation", "reservoir_power", "electrode_power",
                      "energy_balance", "electrode_current", "capacitor_energy", "membrane_heat", "junction_heat",
                      "face_heat")
# the ledgers of the bound ratios, in their column order (CELLS_RATIO_*)
LEDGER_RATIOS = ("local", "total", "global_charge", "water", "outside", "charge", "electrode",
                 "energy", "reservoir_work", "electrode_work", "metabolic_work", "bath_water_work")


class StationaryOptions(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("max_newton", ctypes.c_int32),
                ("max_cycles", ctypes.c_int32), ("linear_rtol", ctypes.c_double),
                ("fp32_inner", ctypes.c_int32)]


class StationaryReport(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("newton_iterations", ctypes.c_int32),
                ("krylov_steps", ctypes.c_int64), ("worst_step", ctypes.c_double)]


class CellsError(RuntimeError):
    """A status other than CELLS_OK; .code is the status."""

    def __init__(self, code, message):
        super().__init__(f"cells_hip error {code}: {message}")
        self.code = code


class WorldStruct(ctypes.Structure):
    _fields_ = [
        ("S", ctypes.c_int32), ("W", ctypes.c_int32), ("water", ctypes.c_int32),
        ("calcium", ctypes.c_int32), ("ncell", ctypes.c_int64), ("nedge", ctypes.c_int64),
        ("cell_off", _i64p), ("edge_off", _i64p),
        ("valence", ctypes.c_double * 4), ("bath_c", ctypes.c_double * 4),
        ("temperature", ctypes.c_double), ("bath_V", ctypes.c_double),
        ("atp", ctypes.c_double), ("adp", ctypes.c_double), ("pi", ctypes.c_double),
        ("km_na", ctypes.c_double), ("km_k", ctypes.c_double), ("km_atp", ctypes.c_double),
        ("dg0", ctypes.c_double), ("ca_km_ca", ctypes.c_double), ("ca_km_atp", ctypes.c_double),
        ("ca_dg0", ctypes.c_double), ("gate_threshold", ctypes.c_double),
        ("gate_minimum", ctypes.c_double), ("water_K", ctypes.c_double),
        ("water_P0", ctypes.c_double), ("water_vmin", ctypes.c_double),
        ("water_vmax", ctypes.c_double), ("bath_pressure", ctypes.c_double),
        ("vol0", _f64p), ("cap", _f64p), ("pam", _f64p), ("nak_cap", _f64p), ("ca_cap", _f64p),
        ("n0", _f64p), ("q0", _f64p), ("lpa", _f64p), ("imperm", _f64p),
        ("snd", _i64p), ("rcv", _i64p), ("paj", _f64p),
        ("inc_ptr", _i64p), ("inc_edge", _i64p), ("inc_out", _i32p),
    ]


class FieldGridStruct(ctypes.Structure):
    _fields_ = [("side", ctypes.c_int32), ("pad", ctypes.c_int32), ("cell_node", _i32p),
                ("node_volume", ctypes.c_double), ("face_factor", ctypes.c_double * 4),
                ("node_n0", _f64p), ("node_q0", _f64p)]


class ControlStruct(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("membrane", "junction", "pump", "calcium_pump", "water")]


_lib = None


def load():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with cells_amd.build.build() (hipcc, gfx950); "
            "the HIP step has no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    lib.cells_last_error.restype = ctypes.c_char_p
    lib.cells_state_size.restype = ctypes.c_int64
    lib.cells_factor_size.restype = ctypes.c_int64
    lib.cells_ensemble_create.argtypes = [ctypes.POINTER(WorldStruct), ctypes.c_int32,
                                          ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    lib.cells_ensemble_destroy.argtypes = [ctypes.c_void_p]
    lib.cells_state_size.argtypes = [ctypes.c_void_p]
    lib.cells_factor_size.argtypes = [ctypes.c_void_p]
    lib.cells_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ControlStruct),
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
    lib.cells_rates.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.POINTER(ControlStruct), ctypes.c_void_p, ctypes.c_void_p]
    lib.cells_energy_rates.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(ControlStruct), ctypes.c_void_p]
    lib.cells_factor.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.cells_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p]
    lib.cells_kernel_times.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.POINTER(ControlStruct),
                                       ctypes.c_int32, ctypes.c_int32, _f64p, ctypes.c_char_p,
                                       _i32p]
    lib.cells_device_checks.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_void_p, ctypes.c_void_p]
    vp = ctypes.c_void_p
    lib.cells_step_vjp.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.POINTER(ControlStruct),
                                   vp, vp, vp, vp, vp]
    lib.cells_step_vjp_noalloc.argtypes = lib.cells_step_vjp.argtypes
    lib.cells_vmem_loss.argtypes = [vp, vp, vp, vp, ctypes.c_double, vp, vp]
    lib.cells_tape_reserve.argtypes = [vp, ctypes.c_int32, _i64p]
    lib.cells_step_checks.argtypes = [vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int64]
    lib.cells_step_check_bits.argtypes = [vp, vp, vp]
    lib.cells_step_bounds.argtypes = [vp, _f64p]
    lib.cells_set_stage_tolerance.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.cells_stationary.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ControlStruct),
                                     ctypes.POINTER(StationaryOptions), vp, vp, vp,
                                     ctypes.POINTER(StationaryReport)]
    lib.cells_stationary_residual.argtypes = [vp, vp, vp, ctypes.POINTER(ControlStruct), vp, vp]
    lib.cells_stationary_jvp.argtypes = [vp, vp, vp, ctypes.POINTER(ControlStruct), vp, vp]
    lib.cells_step_record.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.POINTER(ControlStruct),
                                      ctypes.c_int32, vp]
    lib.cells_step_vjp_recorded.argtypes = [vp, vp, ctypes.c_int32, vp, ctypes.c_double,
                                            ctypes.POINTER(ControlStruct), vp, vp, vp, vp, vp]
    lib.cells_set_parameters.argtypes = [vp, vp, vp, vp, vp]
    lib.cells_plan_stats.argtypes = [vp, _i64p, _f64p, _i32p, _i32p]
    lib.cells_world_positions.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i64p, _i64p, _i64p]
    lib.cells_world_factor_positions.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, _i32p, ctypes.c_int64,
                                                 _i64p]
    lib.cells_minimum_degree.argtypes = [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i32p]
    _u8p = ctypes.POINTER(ctypes.c_uint8)
    lib.cells_capacitor_topology.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _u8p, _u8p, _u8p,
                                             _u8p, _u8p, ctypes.c_int64, _i32p, _i32p, _u8p, ctypes.c_int64, _i32p,
                                             _i32p, _i32p, _u8p, _u8p, _i32p]
    lib.cells_product_plan_size.argtypes = [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i64p]
    lib.cells_product_plan.argtypes = [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i64p, _i32p, _i64p]
    lib.cells_product.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64]
    lib.cells_bind.argtypes = [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i32p,
                               ctypes.POINTER(ctypes.c_void_p)]
    lib.cells_symbolic_sizes.argtypes = [ctypes.c_void_p, _i64p, _i32p]
    lib.cells_symbolic_export.argtypes = [ctypes.c_void_p, _i32p, _i32p, _i64p, _i32p, _i32p,
                                          _i32p, _i32p]
    lib.cells_symbolic_free.argtypes = [ctypes.c_void_p]
    lib.cells_set_plan_options.argtypes = [ctypes.c_int32] * 6
    lib.cells_get_plan_options.argtypes = [ctypes.c_void_p] * 6
    lib.cells_get_plan_options.argtypes = [_i32p] * 4
    lib.cells_relaxed_padding.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _i32p, ctypes.c_int64,
                                          _i64p]
    lib.cells_field_create.argtypes = [ctypes.POINTER(WorldStruct), ctypes.POINTER(FieldGridStruct),
                                       ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
    lib.cells_field_destroy.argtypes = [vp]
    lib.cells_field_state_size.argtypes = [vp]
    lib.cells_field_state_size.restype = ctypes.c_int64
    lib.cells_field_step.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_int32,
                                     ctypes.POINTER(ControlStruct), vp, vp]
    lib.cells_field_refresh.argtypes = [vp]
    lib.cells_field_set_start.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.cells_field_potentials.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.POINTER(ControlStruct), vp, vp, vp]
    lib.cells_field_energy_rates.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.POINTER(ControlStruct), vp, vp]
    lib.cells_field_plan_stats.argtypes = [vp, _i64p, _f64p, _i32p, _i32p, _i32p]
    lib.cells_field_kernel_times.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.POINTER(ControlStruct),
                                             ctypes.c_int32, ctypes.c_int32, _f64p, ctypes.c_char_p, _i32p]
    _lib = lib
    return lib


def check(rc):
    if rc != CELLS_OK:
        raise CellsError(rc, load().cells_last_error().decode())


def ptr(a, t):
    return a.ctypes.data_as(t)


def world_struct(p):
    """The cells_world view of a cells_amd.world.Packed (arrays kept alive on the struct)."""
    c = p.config
    keep = {}

    def arr(name, dtype):
        a = np.ascontiguousarray(getattr(p, name), dtype=dtype)
        keep[name] = a
        return a

    w = WorldStruct()
    w.S, w.W, w.water, w.calcium = p.S, p.W, int(p.water), int(p.calcium)
    w.ncell, w.nedge = p.ncell, p.nedge
    w.cell_off = ptr(arr("cell_off", np.int64), _i64p)
    w.edge_off = ptr(arr("edge_off", np.int64), _i64p)
    for s in range(p.S):
        w.valence[s] = p.valence[s]
        w.bath_c[s] = p.bath_c[s]
    w.temperature, w.bath_V = c.temperature, c.bath_V
    for k in ("atp", "adp", "pi", "km_na", "km_k", "km_atp", "dg0", "ca_km_ca", "ca_km_atp",
              "ca_dg0", "gate_threshold", "gate_minimum", "water_K", "water_P0", "water_vmin",
              "water_vmax", "bath_pressure"):
        setattr(w, k, getattr(c, k))
    for k in ("vol0", "cap", "pam", "nak_cap", "ca_cap", "n0", "q0", "lpa", "imperm", "paj"):
        setattr(w, k, ptr(arr(k, np.float64), _f64p))
    for k in ("snd", "rcv", "inc_ptr", "inc_edge"):
        setattr(w, k, ptr(arr(k, np.int64), _i64p))
    w.inc_out = ptr(arr("inc_out", np.int32), _i32p)
    w._keep = keep
    return w


def field_grid_struct(fp):
    """The cells_field_grid view of a cells_amd.field.FieldPacked (arrays kept alive on it)."""
    keep = [np.ascontiguousarray(fp.cell_node, dtype=np.int32),
            np.ascontiguousarray(fp.node_n0, dtype=np.float64),
            np.ascontiguousarray(fp.node_q0, dtype=np.float64)]
    g = FieldGridStruct()
    g.side = fp.side
    g.cell_node = ptr(keep[0], _i32p)
    g.node_volume = fp.node_volume
    for s in range(fp.S):
        g.face_factor[s] = fp.face_factor[s]
    g.node_n0 = ptr(keep[1], _f64p)
    g.node_q0 = ptr(keep[2], _f64p)
    g._keep = keep
    return g
