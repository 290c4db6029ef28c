"""Python mirror of the calcium_sparks MI355X path (ctypes over the C ABI).

The reference program is C; this module exists so tests and the benchmark
can drive the same C ABI a C caller uses (include/csparks.h, cs_host.h):

* ``Geometry`` / ``SpeciesModel`` / ``Options`` hold what the reference
  reads from its geometry file, species file and command line
  (geometry.c:92-141, species.c:103-258, options.c:502-542);
* ``Simulation`` builds the problem with libcshost and steps it on the GPU
  with libcsparks: ``step`` is the reference loop body
  (calcium_sparks.c:196-213), ``species`` reads a field back in the
  reference layout, ``stats`` gives the per-domain min/max/sum of
  Species_output_scalar_data (species.c:2785-2863).

There is no CPU fallback: constructing a Simulation without the built HIP
library, or without a GPU, raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

MAX_DOMAINS = 5
MAX_SPECIES = 8
MAX_BOUNDARIES = 8
MAX_NAMED = 8
MAX_OPEN = 64
NAME_LEN = 32
MAX_FLUXES = 4

PREC_F64, PREC_F32 = 0, 1
RNG_PHILOX, RNG_MT = 0, 1
BOUNDARY_DENSITY, BOUNDARY_DISCRETE = 0, 1
FLUX_RYR, FLUX_SERCA = 0, 1

Name = C.c_char * NAME_LEN


class CGeometry(C.Structure):
    _fields_ = [
        ("N", C.c_uint64 * 3),
        ("h", C.c_double),
        ("global_size", C.c_double * 3),
        ("num_domains", C.c_uint32),
        ("domain_names", Name * MAX_DOMAINS),
        ("domain_indices", C.c_uint8 * MAX_DOMAINS),
        ("domain_connections", C.c_uint8 * (MAX_DOMAINS * MAX_DOMAINS)),
        ("voxels", C.POINTER(C.c_uint8)),
        ("num_boundaries", C.c_uint32),
        ("boundary_names", Name * MAX_BOUNDARIES),
        ("boundary_types", C.c_uint32 * MAX_BOUNDARIES),
        ("boundary_sizes", C.c_uint64 * MAX_BOUNDARIES),
        ("boundaries", C.POINTER(C.c_uint16) * MAX_BOUNDARIES),
    ]


class CChannelModel(C.Structure):
    _fields_ = [
        ("Kd_open", C.c_double), ("k_min_open", C.c_double), ("k_max_open", C.c_double),
        ("n_open", C.c_double), ("Kd_close", C.c_double), ("k_min_close", C.c_double),
        ("k_max_close", C.c_double), ("n_close", C.c_double), ("i_unitary", C.c_double),
        ("boundary", Name), ("species", Name),
    ]


class CSercaModel(C.Structure):
    _fields_ = [
        ("v_cyt_to_A_sr", C.c_double), ("density", C.c_double), ("scale", C.c_double),
        ("boundary", Name), ("species", Name),
    ]


class CSpeciesModel(C.Structure):
    _fields_ = [
        ("num_domains", C.c_uint32), ("num_species", C.c_uint32),
        ("domain_names", Name * MAX_DOMAINS),
        ("species_names", Name * MAX_SPECIES),
        ("num_diffusive", C.c_uint32 * MAX_DOMAINS),
        ("diffusive", (C.c_uint8 * MAX_SPECIES) * MAX_DOMAINS),
        ("sigma", (C.c_double * MAX_SPECIES) * MAX_DOMAINS),
        ("init", (C.c_double * MAX_SPECIES) * MAX_DOMAINS),
        ("num_buffers", C.c_uint32 * MAX_DOMAINS),
        ("k_on", (C.c_double * MAX_SPECIES) * MAX_DOMAINS),
        ("k_off", (C.c_double * MAX_SPECIES) * MAX_DOMAINS),
        ("tot", (C.c_double * MAX_SPECIES) * MAX_DOMAINS),
        ("buffer_species", ((C.c_uint8 * 2) * MAX_SPECIES) * MAX_DOMAINS),
        ("use_ryr", C.c_uint32), ("use_p_ryr", C.c_uint32), ("use_serca", C.c_uint32),
        ("ryr", CChannelModel), ("p_ryr", CChannelModel), ("serca", CSercaModel),
        ("conv", C.c_double * 3),
        ("num_fixed", C.c_uint32),
        ("fixed", C.c_uint8 * (2 * MAX_DOMAINS * MAX_SPECIES)),
    ]


class COpenEntry(C.Structure):
    _fields_ = [("name", Name), ("count", C.c_uint32), ("idx", C.c_int32 * MAX_OPEN)]


class CCloseEntry(C.Structure):
    _fields_ = [("name", Name), ("time", C.c_double)]


class COptions(C.Structure):
    _fields_ = [
        ("tstop", C.c_double), ("h", C.c_double), ("dt", C.c_double), ("dt_save", C.c_double),
        ("dt_update_react", C.c_double), ("dt_update_stoch", C.c_double), ("tabort", C.c_double),
        ("force_dirichlet", C.c_int32), ("seed_given", C.c_int32), ("seed", C.c_uint32),
        ("num_open", C.c_uint32), ("open", COpenEntry * MAX_NAMED),
        ("num_t_close", C.c_uint32), ("t_close", CCloseEntry * MAX_NAMED),
        ("precision", C.c_uint32), ("rng", C.c_uint32),
    ]


class CFluxDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("species", C.c_uint32), ("substep", C.c_uint32),
        ("inner", C.c_uint32), ("discrete", C.c_uint32), ("num_pairs", C.c_uint32),
        ("pairs", C.POINTER(C.c_uint32)), ("params", C.c_double * 10),
        ("init_states", C.POINTER(C.c_uint8)),
    ]


class CProblem(C.Structure):
    _fields_ = [
        ("n", C.c_uint32 * 3), ("num_domains", C.c_uint32), ("num_species", C.c_uint32),
        ("precision", C.c_uint32), ("rng", C.c_uint32), ("seed", C.c_uint64),
        ("instance", C.c_uint64), ("h", C.c_double), ("dt", C.c_double), ("dV", C.c_double),
        ("domain_ids", C.POINTER(C.c_uint8)),
        ("init", (C.c_double * MAX_SPECIES) * MAX_DOMAINS),
        ("substeps", C.c_uint32 * MAX_SPECIES),
        ("coef", (C.c_double * (MAX_DOMAINS * MAX_DOMAINS)) * MAX_SPECIES),
        ("num_buffers", C.c_uint32), ("reaction_substep", C.c_uint32), ("dt_react", C.c_double),
        ("bsp0", (C.c_uint8 * MAX_DOMAINS) * MAX_SPECIES),
        ("bsp1", (C.c_uint8 * MAX_DOMAINS) * MAX_SPECIES),
        ("k_on", (C.c_double * MAX_DOMAINS) * MAX_SPECIES),
        ("k_off", (C.c_double * MAX_DOMAINS) * MAX_SPECIES),
        ("tot", (C.c_double * MAX_DOMAINS) * MAX_SPECIES),
        ("num_fluxes", C.c_uint32), ("fluxes", CFluxDesc * MAX_FLUXES),
        ("stochastic_substep", C.c_uint32), ("force_dirichlet", C.c_uint32),
        ("fixed", (C.c_uint8 * MAX_SPECIES) * MAX_DOMAINS),
    ]


class CMT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_uint32)]


class CDiscreteInfo(C.Structure):
    _fields_ = [("name", Name), ("boundary", C.c_uint32), ("flux", C.c_int32), ("size", C.c_uint32)]


class CSetup(C.Structure):
    _fields_ = [
        ("problem", CProblem), ("subdivisions", C.c_uint32), ("h_geom", C.c_double),
        ("N", C.c_uint64 * 3), ("num_domains", C.c_uint32), ("num_species", C.c_uint32),
        ("domain_names", Name * MAX_DOMAINS), ("species_names", Name * MAX_SPECIES),
        ("volumes", C.c_double * MAX_DOMAINS), ("domain_voxels", C.c_int64 * MAX_DOMAINS),
        ("num_discrete", C.c_uint32), ("discrete", CDiscreteInfo * MAX_BOUNDARIES),
        ("num_diffusive_species", C.c_uint32), ("all_diffusive", C.c_uint8 * MAX_SPECIES),
        ("has_ryr_boundary", C.c_uint32), ("has_p_ryr_boundary", C.c_uint32),
        ("ryr_discrete", C.c_int32), ("p_ryr_discrete", C.c_int32),
        ("mt", CMT), ("seed", C.c_uint32),
    ]


class CLaunchConfig(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("zchunk", C.c_uint32), ("xcd_remap", C.c_uint32),
                ("prefetch", C.c_uint32), ("nontemporal", C.c_uint32), ("lds_tile", C.c_uint32),
                ("linear", C.c_uint32), ("yblock", C.c_uint32), ("fuse", C.c_uint32),
                ("fuse_zchunk", C.c_uint32), ("fuse_xcd", C.c_uint32), ("fuse_fixed", C.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class CBatchLaunch(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("zchunk", C.c_uint32), ("prefetch", C.c_uint32),
                ("nontemporal", C.c_uint32), ("persistent", C.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class CAbortState(C.Structure):
    _fields_ = [("tabort", C.c_double), ("last_closed", C.c_double),
                ("all_closed", C.c_int), ("active", C.c_int)]


# --------------------------------------------------------------- libraries
_libs: Dict[str, C.CDLL] = {}


def _load(name: str) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run hpc-ca-phospattern_amd/build.py")
    lib = C.CDLL(path)
    _libs[name] = lib
    return lib


def host_lib() -> C.CDLL:
    lib = _load("libcshost.so")
    if not getattr(lib, "_cs_typed", False):
        lib.cs_setup_build.restype = C.c_int
        lib.cs_setup_build.argtypes = [C.POINTER(CGeometry), C.POINTER(CSpeciesModel),
                                       C.POINTER(COptions), C.POINTER(C.POINTER(CSetup)),
                                       C.c_char_p, C.c_size_t]
        lib.cs_setup_free.argtypes = [C.POINTER(CSetup)]
        lib.cs_setup_free.restype = None
        lib.cs_setup_mt_uniforms.argtypes = [C.POINTER(CSetup), C.c_uint64, C.c_uint64,
                                             C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        lib.cs_options_defaults.argtypes = [C.POINTER(COptions)]
        lib.cs_options_defaults.restype = None
        lib.cs_mt_seed.argtypes = [C.POINTER(CMT), C.c_uint32]
        lib.cs_mt_seed.restype = None
        lib.cs_mt_u32.argtypes = [C.POINTER(CMT)]
        lib.cs_mt_u32.restype = C.c_uint32
        lib.cs_mt_ldrand.argtypes = [C.POINTER(CMT)]
        lib.cs_mt_ldrand.restype = C.c_double
        lib.cs_save_interval.argtypes = [C.c_double, C.c_double]
        lib.cs_save_interval.restype = C.c_uint64
        lib.cs_count_steps.argtypes = [C.c_double, C.c_double]
        lib.cs_count_steps.restype = C.c_uint64
        lib.cs_abort_init.argtypes = [C.POINTER(CAbortState), C.c_double, C.c_int]
        lib.cs_abort_init.restype = None
        lib.cs_abort_snapshot.argtypes = [C.POINTER(CAbortState), C.c_int]
        lib.cs_abort_snapshot.restype = None
        lib.cs_abort_check.argtypes = [C.POINTER(CAbortState), C.c_double]
        lib.cs_abort_check.restype = C.c_int
        lib.cs_host_sizeof.argtypes = [C.c_int]
        lib.cs_host_sizeof.restype = C.c_size_t
        lib._cs_typed = True
    return lib


def device_lib() -> C.CDLL:
    lib = _load("libcsparks.so")
    if not getattr(lib, "_cs_typed", False):
        P = C.c_void_p
        lib.cs_create.argtypes = [C.POINTER(CProblem), C.POINTER(P)]
        lib.cs_destroy.argtypes = [P]
        lib.cs_reset.argtypes = [P]
        lib.cs_upload_species.argtypes = [P, C.c_uint32, C.POINTER(C.c_double)]
        lib.cs_download_species.argtypes = [P, C.c_uint32, C.POINTER(C.c_double)]
        lib.cs_step.argtypes = [P, C.c_uint64, C.c_uint64]
        lib.cs_step_uniforms.argtypes = [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_double), C.c_uint64]
        lib.cs_uniforms_needed.argtypes = [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        lib.cs_synchronize.argtypes = [P]
        lib.cs_reduce_stats.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.cs_get_channel_recruited.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint8)]
        lib.cs_linescan.argtypes = [P, C.c_uint32, C.c_uint32] + [C.POINTER(C.c_double)] * 3 + \
            [C.c_uint32, C.POINTER(C.c_double)]
        lib.cs_get_channel_states.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint8)]
        lib.cs_set_channel_states.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint8)]
        lib.cs_time_steps.argtypes = [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.cs_time_passes.argtypes = [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        lib.cs_fused_available.argtypes = [P, C.c_char_p, C.c_size_t]
        lib.cs_device_info.argtypes = [C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
        lib.cs_ryr_flux.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)]
        lib.cs_ryr_flux.restype = C.c_double
        lib.cs_serca_flux.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)]
        lib.cs_serca_flux.restype = C.c_double
        lib.cs_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.cs_philox4x32_10.restype = None
        lib.cs_get_launch_config.argtypes = [P, C.POINTER(CLaunchConfig)]
        lib.cs_set_launch_config.argtypes = [P, C.POINTER(CLaunchConfig)]
        lib.cs_autotune.argtypes = [P, C.c_uint64, C.c_uint32, C.POINTER(CLaunchConfig), C.POINTER(C.c_double)]
        lib.cs_calibration_copy.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        lib.cs_last_error.restype = C.c_char_p
        pd = C.POINTER(C.c_double)
        lib.cs_batch_create.argtypes = [C.POINTER(CProblem), C.c_uint32, C.POINTER(P)]
        lib.cs_batch_destroy.argtypes = [P]
        lib.cs_batch_reset.argtypes = [P]
        lib.cs_batch_info.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        lib.cs_batch_upload_species.argtypes = [P, C.c_uint32, C.c_uint32, pd]
        lib.cs_batch_download_species.argtypes = [P, C.c_uint32, C.c_uint32, pd]
        lib.cs_batch_step.argtypes = [P, C.c_uint64, C.c_uint64]
        lib.cs_batch_step_uniforms.argtypes = [P, C.c_uint64, C.c_uint64, pd, C.c_uint64]
        lib.cs_batch_uniforms_needed.argtypes = [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        lib.cs_batch_synchronize.argtypes = [P]
        lib.cs_batch_reduce_stats.argtypes = [P, pd, pd, pd]
        lib.cs_batch_linescan.argtypes = [P, C.c_uint32, C.c_uint32, pd, C.c_uint32, pd]
        lib.cs_batch_get_channel_states.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]
        lib.cs_batch_get_channel_recruited.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]
        lib.cs_batch_channel_summary.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                                 C.POINTER(C.c_uint32)]
        lib.cs_batch_set_active.argtypes = [P, C.POINTER(C.c_uint8)]
        lib.cs_batch_get_launch.argtypes = [P, C.POINTER(CBatchLaunch)]
        lib.cs_batch_set_launch.argtypes = [P, C.POINTER(CBatchLaunch)]
        lib.cs_batch_autotune.argtypes = [P, C.c_uint64, C.c_uint32, C.POINTER(CBatchLaunch), pd]
        lib.cs_batch_set_graph.argtypes = [P, C.c_uint32]
        lib.cs_batch_step_period.argtypes = [P, C.POINTER(C.c_uint64)]
        lib.cs_batch_time_steps.argtypes = [P, C.c_uint64, C.c_uint64, pd, pd]
        pf = C.POINTER(C.c_float)
        lib.cs_batch_set_gating.argtypes = [P, C.c_uint32, P]
        lib.cs_batch_copy_channel_states.argtypes = [P, P]
        lib.cs_batch_channel_count.argtypes = [P, C.POINTER(C.c_uint32)]
        lib.cs_batch_closure_features.argtypes = [P, pf]
        lib.cs_batch_channel_ca.argtypes = [P, pd]
        lib.cs_batch_closure_predictions.argtypes = [P, pd]
        lib.cs_batch_channel_sr_ca.argtypes = [P, pd]
        lib.cs_batch_channel_open.argtypes = [P, C.POINTER(C.c_uint8)]
        lib.cs_batch_set_flux_closure.argtypes = [P, P]
        lib.cs_batch_flux_predictions.argtypes = [P, pd]
        lib.cs_batch_monitor.argtypes = [P, C.POINTER(C.c_uint32)]
        lib.cs_mlp_create.argtypes = [pf, C.POINTER(P)]
        lib.cs_mlp_destroy.argtypes = [P]
        lib.cs_mlp_get_params.argtypes = [P, pf]
        lib.cs_mlp_set_params.argtypes = [P, pf]
        lib.cs_mlp_set_normalisation.argtypes = [P, pf, pf, pf, pf, C.c_float, C.c_float, C.c_float]
        lib.cs_mlp_forward.argtypes = [P, pf, C.c_uint64, pf]
        lib.cs_mlp_grad.argtypes = [P, pf, pf, C.c_uint64, pf, pd]
        lib.cs_mlp_set_data.argtypes = [P, pf, pf, C.c_uint64, C.c_uint64]
        lib.cs_mlp_train.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, pd]
        lib._cs_typed = True
    return lib


def io_lib() -> C.CDLL:
    lib = _load("libcsio.so")
    if not getattr(lib, "_cs_typed", False):
        lib.cs_io_read_geometry.argtypes = [C.c_char_p, C.POINTER(CGeometry), C.c_char_p, C.c_size_t]
        lib.cs_io_free_geometry.argtypes = [C.POINTER(CGeometry)]
        lib.cs_io_free_geometry.restype = None
        lib.cs_io_read_species.argtypes = [C.c_char_p, C.POINTER(CSpeciesModel), C.c_char_p, C.c_size_t]
        lib.cs_io_write_geometry.argtypes = [C.c_char_p, C.POINTER(CGeometry), C.c_char_p, C.c_size_t]
        lib.cs_io_write_species.argtypes = [C.c_char_p, C.POINTER(CSpeciesModel), C.c_char_p, C.c_size_t]
        lib.cs_io_read_dataset.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_uint64,
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        lib.cs_io_read_attr.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_uint64,
                                        C.POINTER(C.c_uint64)]
        lib.cs_io_read_attr_str.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
        lib._cs_typed = True
    return lib


class CSError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != 0:
        msg = device_lib().cs_last_error().decode(errors="replace")
        raise CSError(f"{what} failed ({rc}): {msg}")


# ------------------------------------------------------------ input model
@dataclass
class Boundary:
    name: str
    type: str                      # "density" | "discrete"
    pairs: np.ndarray              # uint16 [n, 6] geometry voxels (from, to)


@dataclass
class Geometry:
    """Content of a reference geometry file (geometry.c:92-141, :487-1011)."""
    voxels: np.ndarray             # uint8 [N0, N1, N2], z fastest
    h: float
    domain_names: List[str]
    domain_connections: np.ndarray  # uint8 [nd, nd]
    boundaries: List[Boundary] = field(default_factory=list)
    global_size: Optional[Tuple[float, float, float]] = None

    def boundary(self, name: str) -> Boundary:
        for b in self.boundaries:
            if b.name == name:
                return b
        raise KeyError(name)


@dataclass
class ChannelModel:
    Kd_open: float
    k_min_open: float
    k_max_open: float
    n_open: float
    Kd_close: float
    k_min_close: float
    k_max_close: float
    n_close: float
    i_unitary: float
    boundary: str
    species: str = "Ca"


@dataclass
class SercaModel:
    v_cyt_to_A_sr: float
    density: float
    scale: float
    boundary: str
    species: str = "Ca"


@dataclass
class Buffer:
    bound: int                     # species index of the buffer
    free: int                      # species index it binds (Ca)
    k_on: float
    k_off: float
    tot: float


@dataclass
class SpeciesModel:
    """Content of a reference species file (species.c:103-258)."""
    domain_names: List[str]
    species_names: List[str]
    diffusive: List[List[int]]     # per domain: species indices
    sigma: List[List[float]]       # per domain: nm^2/ms, same order
    init: np.ndarray               # [nd, ns] uM
    buffers: List[List[Buffer]]    # per domain
    ryr: Optional[ChannelModel] = None
    p_ryr: Optional[ChannelModel] = None
    serca: Optional[SercaModel] = None
    conv: Tuple[float, float, float] = (28853.9, 28853.9, 115416.0)
    fixed: List[Tuple[int, int]] = field(default_factory=list)


def gauss(x, sigma):
    """Gaussian weight of the linescan (gauss, utils.c:209-213)."""
    return np.sqrt(0.5) * (2.0 / np.sqrt(np.pi)) * (1.0 / np.sqrt(sigma)) / 2 * np.exp(-np.asarray(x) * x / 2 / sigma)


@dataclass
class Options:
    """Command-line options of the reference (options.c:502-542)."""
    tstop: float = 10.0
    h: float = 6.0
    dt: float = -1.0
    dt_save: float = 0.5
    dt_update_react: float = -1.0
    dt_update_stoch: float = -1.0
    tabort: float = 0.0
    force_dirichlet: bool = False
    seed: Optional[int] = None
    open: Dict[str, Sequence[int]] = field(default_factory=dict)
    t_close: Dict[str, float] = field(default_factory=dict)
    precision: str = "f64"         # "f64" | "f32"
    rng: str = "philox"            # "philox" | "mt"


def _setname(dst, s: str):
    b = s.encode()
    if len(b) >= NAME_LEN:
        raise ValueError(f"name too long: {s}")
    dst.value = b


def _setfield(struct, fname: str, s: str):
    b = s.encode()
    if len(b) >= NAME_LEN:
        raise ValueError(f"name too long: {s}")
    setattr(struct, fname, b)


class _Holder:
    """Keeps numpy buffers alive while C structs point into them."""

    def __init__(self):
        self.keep: List[object] = []


def geometry_to_c(g: Geometry, hold: _Holder) -> CGeometry:
    cg = CGeometry()
    vox = np.ascontiguousarray(g.voxels, dtype=np.uint8)
    hold.keep.append(vox)
    for a in range(3):
        cg.N[a] = vox.shape[a]
        cg.global_size[a] = (g.global_size[a] if g.global_size else vox.shape[a] * g.h)
    cg.h = g.h
    nd = len(g.domain_names)
    if nd > MAX_DOMAINS:
        raise ValueError("too many domains")
    cg.num_domains = nd
    for i, n in enumerate(g.domain_names):
        _setname(cg.domain_names[i], n)
        cg.domain_indices[i] = i
    conn = np.asarray(g.domain_connections, dtype=np.uint8).reshape(nd, nd)
    for i in range(nd):
        for j in range(nd):
            cg.domain_connections[i * nd + j] = int(conn[i, j])
    cg.voxels = vox.ctypes.data_as(C.POINTER(C.c_uint8))
    if len(g.boundaries) > MAX_BOUNDARIES:
        raise ValueError("too many boundaries")
    cg.num_boundaries = len(g.boundaries)
    for i, b in enumerate(g.boundaries):
        arr = np.ascontiguousarray(b.pairs, dtype=np.uint16).reshape(-1, 6)
        hold.keep.append(arr)
        _setname(cg.boundary_names[i], b.name)
        cg.boundary_types[i] = BOUNDARY_DISCRETE if b.type == "discrete" else BOUNDARY_DENSITY
        cg.boundary_sizes[i] = arr.shape[0]
        cg.boundaries[i] = arr.ctypes.data_as(C.POINTER(C.c_uint16))
    return cg


def _channel_to_c(dst: CChannelModel, m: ChannelModel):
    for k in ("Kd_open", "k_min_open", "k_max_open", "n_open", "Kd_close", "k_min_close",
              "k_max_close", "n_close", "i_unitary"):
        setattr(dst, k, float(getattr(m, k)))
    _setfield(dst, "boundary", m.boundary)
    _setfield(dst, "species", m.species)


def model_to_c(m: SpeciesModel) -> CSpeciesModel:
    cm = CSpeciesModel()
    nd, ns = len(m.domain_names), len(m.species_names)
    if nd > MAX_DOMAINS or ns > MAX_SPECIES:
        raise ValueError("too many domains or species")
    cm.num_domains, cm.num_species = nd, ns
    for i, n in enumerate(m.domain_names):
        _setname(cm.domain_names[i], n)
    for i, n in enumerate(m.species_names):
        _setname(cm.species_names[i], n)
    init = np.asarray(m.init, dtype=np.float64).reshape(nd, ns)
    for d in range(nd):
        cm.num_diffusive[d] = len(m.diffusive[d])
        for k, (sp, sg) in enumerate(zip(m.diffusive[d], m.sigma[d])):
            cm.diffusive[d][k] = sp
            cm.sigma[d][k] = sg
        for s in range(ns):
            cm.init[d][s] = init[d, s]
        cm.num_buffers[d] = len(m.buffers[d])
        for k, b in enumerate(m.buffers[d]):
            cm.k_on[d][k], cm.k_off[d][k], cm.tot[d][k] = b.k_on, b.k_off, b.tot
            cm.buffer_species[d][k][0] = b.bound
            cm.buffer_species[d][k][1] = b.free
    if m.ryr is not None:
        cm.use_ryr = 1
        _channel_to_c(cm.ryr, m.ryr)
    if m.p_ryr is not None:
        cm.use_p_ryr = 1
        _channel_to_c(cm.p_ryr, m.p_ryr)
    if m.serca is not None:
        cm.use_serca = 1
        cm.serca.v_cyt_to_A_sr = m.serca.v_cyt_to_A_sr
        cm.serca.density = m.serca.density
        cm.serca.scale = m.serca.scale
        _setfield(cm.serca, "boundary", m.serca.boundary)
        # the reference takes the SERCA species from /ryr/species
        _setfield(cm.serca, "species", m.ryr.species if m.ryr is not None else m.serca.species)
    for a in range(3):
        cm.conv[a] = m.conv[a]
    cm.num_fixed = len(m.fixed)
    for i, (d, s) in enumerate(m.fixed):
        cm.fixed[2 * i], cm.fixed[2 * i + 1] = d, s
    return cm


def options_to_c(o: Options) -> COptions:
    co = COptions()
    host_lib().cs_options_defaults(C.byref(co))
    co.tstop, co.h, co.dt, co.dt_save = o.tstop, o.h, o.dt, o.dt_save
    co.dt_update_react, co.dt_update_stoch, co.tabort = o.dt_update_react, o.dt_update_stoch, o.tabort
    co.force_dirichlet = 1 if o.force_dirichlet else 0
    if o.seed is not None:
        co.seed_given, co.seed = 1, int(o.seed) & 0xFFFFFFFF
    if len(o.open) > MAX_NAMED or len(o.t_close) > MAX_NAMED:
        raise ValueError("too many -O / -C entries")
    co.num_open = len(o.open)
    for i, (name, idx) in enumerate(o.open.items()):
        _setfield(co.open[i], "name", name)
        idx = list(idx)
        if len(idx) > MAX_OPEN:
            raise ValueError("too many open indices")
        co.open[i].count = len(idx)
        for j, v in enumerate(idx):
            co.open[i].idx[j] = int(v)
    co.num_t_close = len(o.t_close)
    for i, (name, t) in enumerate(o.t_close.items()):
        _setfield(co.t_close[i], "name", name)
        co.t_close[i].time = float(t)
    co.precision = PREC_F32 if o.precision == "f32" else PREC_F64
    co.rng = RNG_MT if o.rng == "mt" else RNG_PHILOX
    return co


def _s(b) -> str:
    return b.decode() if isinstance(b, bytes) else b.value.decode()


def geometry_from_c(cg: CGeometry) -> Geometry:
    N = tuple(int(v) for v in cg.N)
    vox = np.ctypeslib.as_array(cg.voxels, shape=(int(np.prod(N)),)).reshape(N).copy()
    nd = int(cg.num_domains)
    conn = np.array([cg.domain_connections[i] for i in range(nd * nd)], dtype=np.uint8).reshape(nd, nd)
    bnds = []
    for i in range(int(cg.num_boundaries)):
        n = int(cg.boundary_sizes[i])
        arr = (np.ctypeslib.as_array(cg.boundaries[i], shape=(n * 6,)).reshape(n, 6).copy()
               if n else np.zeros((0, 6), np.uint16))
        bnds.append(Boundary(_s(cg.boundary_names[i]),
                             "discrete" if cg.boundary_types[i] == BOUNDARY_DISCRETE else "density", arr))
    return Geometry(voxels=vox, h=cg.h, domain_names=[_s(cg.domain_names[i]) for i in range(nd)],
                    domain_connections=conn, boundaries=bnds, global_size=tuple(cg.global_size))


def _channel_from_c(c: CChannelModel) -> ChannelModel:
    return ChannelModel(c.Kd_open, c.k_min_open, c.k_max_open, c.n_open, c.Kd_close, c.k_min_close,
                        c.k_max_close, c.n_close, c.i_unitary, _s(c.boundary), _s(c.species))


def model_from_c(cm: CSpeciesModel) -> SpeciesModel:
    nd, ns = int(cm.num_domains), int(cm.num_species)
    diffusive = [[int(cm.diffusive[d][k]) for k in range(cm.num_diffusive[d])] for d in range(nd)]
    sigma = [[float(cm.sigma[d][k]) for k in range(cm.num_diffusive[d])] for d in range(nd)]
    init = np.array([[cm.init[d][s] for s in range(ns)] for d in range(nd)])
    buffers = [[Buffer(int(cm.buffer_species[d][b][0]), int(cm.buffer_species[d][b][1]), cm.k_on[d][b],
                       cm.k_off[d][b], cm.tot[d][b]) for b in range(cm.num_buffers[d])] for d in range(nd)]
    return SpeciesModel(
        domain_names=[_s(cm.domain_names[d]) for d in range(nd)],
        species_names=[_s(cm.species_names[s]) for s in range(ns)],
        diffusive=diffusive, sigma=sigma, init=init, buffers=buffers,
        ryr=_channel_from_c(cm.ryr) if cm.use_ryr else None,
        p_ryr=_channel_from_c(cm.p_ryr) if cm.use_p_ryr else None,
        serca=SercaModel(cm.serca.v_cyt_to_A_sr, cm.serca.density, cm.serca.scale, _s(cm.serca.boundary),
                         _s(cm.serca.species)) if cm.use_serca else None,
        conv=tuple(cm.conv), fixed=[(int(cm.fixed[2 * i]), int(cm.fixed[2 * i + 1])) for i in range(cm.num_fixed)])


def read_geometry(path: str) -> Geometry:
    """Read a reference geometry file (HDF5)."""
    cg = CGeometry()
    err = C.create_string_buffer(512)
    if io_lib().cs_io_read_geometry(path.encode(), C.byref(cg), err, 512) != 0:
        raise CSError(err.value.decode())
    try:
        return geometry_from_c(cg)
    finally:
        io_lib().cs_io_free_geometry(C.byref(cg))


def read_species(path: str) -> SpeciesModel:
    """Read a reference species file (HDF5)."""
    cm = CSpeciesModel()
    err = C.create_string_buffer(512)
    if io_lib().cs_io_read_species(path.encode(), C.byref(cm), err, 512) != 0:
        raise CSError(err.value.decode())
    return model_from_c(cm)


def read_output_dataset(path: str, dataset: str, max_elems: int = 1 << 28) -> np.ndarray:
    """A float64 dataset of a calcium_sparks output file."""
    out = np.empty(max_elems if max_elems < (1 << 16) else 0, dtype=np.float64)
    dims = (C.c_uint64 * 3)()
    rank = C.c_int()
    # query the size with a zero-length read first
    probe = io_lib().cs_io_read_dataset(path.encode(), dataset.encode(), None, 0, dims, C.byref(rank))
    n = int(np.prod([dims[i] for i in range(rank.value)])) if rank.value else 0
    if rank.value == 0:
        raise CSError(f"no dataset {dataset} in {path}")
    out = np.empty(max(n, 1), dtype=np.float64)
    if io_lib().cs_io_read_dataset(path.encode(), dataset.encode(), out.ctypes.data_as(C.POINTER(C.c_double)),
                                   n, dims, C.byref(rank)) != 0:
        raise CSError(f"cannot read {dataset} from {path}")
    del probe
    return out[:n].reshape([dims[i] for i in range(rank.value)])


def read_output_attr(path: str, obj: str, attr: str) -> np.ndarray:
    out = np.empty(64, dtype=np.float64)
    cnt = C.c_uint64()
    if io_lib().cs_io_read_attr(path.encode(), obj.encode(), attr.encode(),
                                out.ctypes.data_as(C.POINTER(C.c_double)), 64, C.byref(cnt)) != 0:
        raise CSError(f"no attribute {obj}:{attr} in {path}")
    return out[:cnt.value]


def read_output_attr_str(path: str, obj: str, attr: str) -> str:
    buf = C.create_string_buffer(512)
    if io_lib().cs_io_read_attr_str(path.encode(), obj.encode(), attr.encode(), buf, 512) != 0:
        raise CSError(f"no attribute {obj}:{attr} in {path}")
    return buf.value.decode()


def write_geometry(path: str, g: Geometry):
    hold = _Holder()
    cg = geometry_to_c(g, hold)
    err = C.create_string_buffer(512)
    if io_lib().cs_io_write_geometry(path.encode(), C.byref(cg), err, 512) != 0:
        raise CSError(err.value.decode())


def write_species(path: str, m: SpeciesModel):
    cm = model_to_c(m)
    err = C.create_string_buffer(512)
    if io_lib().cs_io_write_species(path.encode(), C.byref(cm), err, 512) != 0:
        raise CSError(err.value.decode())


class Setup:
    """cs_setup_build result: the derived problem plus host metadata."""

    def __init__(self, geom: Geometry, model: SpeciesModel, opts: Options):
        self._hold = _Holder()
        self.cgeom = geometry_to_c(geom, self._hold)
        self.cmodel = model_to_c(model)
        self.copts = options_to_c(opts)
        err = C.create_string_buffer(512)
        ptr = C.POINTER(CSetup)()
        rc = host_lib().cs_setup_build(C.byref(self.cgeom), C.byref(self.cmodel), C.byref(self.copts),
                                       C.byref(ptr), err, 512)
        if rc != 0:
            raise CSError(err.value.decode(errors="replace"))
        self.ptr = ptr
        self.s = ptr.contents

    @property
    def problem(self) -> CProblem:
        return self.s.problem

    @property
    def n(self) -> Tuple[int, int, int]:
        return tuple(int(v) for v in self.s.problem.n)

    @property
    def dt(self) -> float:
        return self.s.problem.dt

    def mt_uniforms(self, first: int, nsteps: int) -> np.ndarray:
        cnt = C.c_uint64()
        host_lib().cs_setup_mt_uniforms(self.ptr, first, nsteps, None, C.byref(cnt))
        out = np.empty(cnt.value, dtype=np.float64)
        host_lib().cs_setup_mt_uniforms(self.ptr, first, nsteps,
                                        out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cnt))
        return out

    def close(self):
        if getattr(self, "ptr", None):
            host_lib().cs_setup_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """A resident calcium_sparks domain on the current GPU."""

    def __init__(self, geom: Geometry, model: SpeciesModel, opts: Options, instance: int = 0):
        self.setup = Setup(geom, model, opts)
        self.setup.s.problem.instance = int(instance)   # Philox key word 1
        self.opts = opts
        self.ctx = C.c_void_p()
        _check(device_lib().cs_create(C.byref(self.setup.s.problem), C.byref(self.ctx)), "cs_create")
        self.time_ind = 1          # next time index (the reference starts at 1)
        self.ns = int(self.setup.s.num_species)
        self.nd = int(self.setup.s.num_domains)

    @property
    def n(self):
        return self.setup.n

    @property
    def dt(self):
        return self.setup.dt

    @property
    def problem(self) -> CProblem:
        return self.setup.problem

    def step(self, nsteps: int = 1):
        lib = device_lib()
        if self.setup.problem.rng == RNG_MT:
            u = self.setup.mt_uniforms(self.time_ind, nsteps)
            _check(lib.cs_step_uniforms(self.ctx, self.time_ind, nsteps,
                                        u.ctypes.data_as(C.POINTER(C.c_double)), len(u)), "cs_step_uniforms")
        else:
            _check(lib.cs_step(self.ctx, self.time_ind, nsteps), "cs_step")
        self.time_ind += nsteps

    def synchronize(self):
        _check(device_lib().cs_synchronize(self.ctx), "cs_synchronize")

    def species(self, s: int) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        _check(device_lib().cs_download_species(self.ctx, s, out.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_download_species")
        return out

    def set_species(self, s: int, values: np.ndarray):
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n)
        _check(device_lib().cs_upload_species(self.ctx, s, arr.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_upload_species")

    def stats(self):
        k = self.nd * self.ns
        mn, mx, tot = (np.empty(k), np.empty(k), np.empty(k))
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        _check(device_lib().cs_reduce_stats(self.ctx, ptr(mn), ptr(mx), ptr(tot)), "cs_reduce_stats")
        return mn.reshape(self.nd, self.ns), mx.reshape(self.nd, self.ns), tot.reshape(self.nd, self.ns)

    def linescan(self, species: int, axis: int, offsets, conv, num_domains: int = 0) -> np.ndarray:
        """Linescan of one species along a reference axis, as written by
        Species_write_convolution_to_file (species.c:1960-2102): the
        transverse Gaussian weights project the field on the device
        (cs_linescan), and the projection is convolved along the axis."""
        n = self.n
        h = self.setup.problem.h
        off = [(n[ax] // 2 + int(offsets[ax])) * h for ax in range(3)]
        w = [np.ascontiguousarray(gauss(off[ax] - np.arange(n[ax]) * h, conv[ax])) for ax in range(3)]
        mask = 0xFFFFFFFF if not num_domains else (0xFF if num_domains >= 8 else (1 << num_domains) - 1)
        proj = np.empty(n[axis])
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        _check(device_lib().cs_linescan(self.ctx, species, axis, ptr(w[0]), ptr(w[1]), ptr(w[2]), mask, ptr(proj)),
               "cs_linescan")
        k = np.arange(n[axis]) * h
        return np.array([float(np.dot(proj, gauss(p - k, conv[axis]))) for p in k])

    def recruited(self, flux: int) -> np.ndarray:
        """Channels of discrete flux `flux` open at some gating step so far
        (cs_get_channel_recruited)."""
        n = self.setup.problem.fluxes[flux].num_pairs
        out = np.zeros(max(n, 1), dtype=np.uint8)
        _check(device_lib().cs_get_channel_recruited(self.ctx, flux, out.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_get_channel_recruited")
        return out[:n]

    def channel_states(self, flux: int) -> np.ndarray:
        n = self.setup.problem.fluxes[flux].num_pairs
        out = np.zeros(max(n, 1), dtype=np.uint8)
        _check(device_lib().cs_get_channel_states(self.ctx, flux, out.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_get_channel_states")
        return out[:n]

    def set_channel_states(self, flux: int, states):
        arr = np.ascontiguousarray(states, dtype=np.uint8)
        _check(device_lib().cs_set_channel_states(self.ctx, flux, arr.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_set_channel_states")

    def launch_config(self) -> dict:
        cfg = CLaunchConfig()
        _check(device_lib().cs_get_launch_config(self.ctx, C.byref(cfg)), "cs_get_launch_config")
        return cfg.as_dict()

    def set_launch_config(self, **kw):
        cfg = CLaunchConfig(**{**self.launch_config(), **kw})
        _check(device_lib().cs_set_launch_config(self.ctx, C.byref(cfg)), "cs_set_launch_config")

    def autotune(self, reps: int = 1):
        """Pick the fastest launch shape of the stencil kernel for this grid
        over the coming steps' substep pattern (the state is not advanced);
        returns (config, mean kernel ms)."""
        cfg, ms = CLaunchConfig(), C.c_double()
        _check(device_lib().cs_autotune(self.ctx, self.time_ind, reps, C.byref(cfg), C.byref(ms)), "cs_autotune")
        return cfg.as_dict(), ms.value

    def time_steps(self, nsteps: int):
        """Time nsteps steps with HIP events; returns (total_ms, mean stencil pass ms)."""
        tot, ker = C.c_double(), C.c_double()
        _check(device_lib().cs_time_steps(self.ctx, self.time_ind, nsteps, C.byref(tot), C.byref(ker)),
               "cs_time_steps")
        self.time_ind += nsteps
        return tot.value, ker.value

    def time_passes(self, nsteps: int):
        """Time nsteps steps; returns (total_ms, mean pass ms, passes, fused passes)."""
        tot, ker, npass, nfused = C.c_double(), C.c_double(), C.c_uint64(), C.c_uint64()
        _check(device_lib().cs_time_passes(self.ctx, self.time_ind, nsteps, C.byref(tot), C.byref(ker),
                                           C.byref(npass), C.byref(nfused)), "cs_time_passes")
        self.time_ind += nsteps
        return tot.value, ker.value, npass.value, nfused.value

    def fused_available(self):
        """(True, '') when the problem admits fused two-step passes, else (False, reason)."""
        why = C.create_string_buffer(256)
        ok = device_lib().cs_fused_available(self.ctx, why, 256)
        if ok < 0:
            _check(ok, "cs_fused_available")
        return ok == 1, why.value.decode()

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            device_lib().cs_destroy(self.ctx)
            self.ctx = C.c_void_p()
        if getattr(self, "setup", None) is not None:
            self.setup.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Realisations of one species model stepped together on the current GPU
    (cs_batch_*): each member is a (geometry, model, options) triple whose
    model and options may differ only in what the geometry, the channel
    layout and the seed decide.  Member i takes the same steps as a
    Simulation of the same triple (its Philox key is the seed its setup
    chose and the problem's instance word)."""

    def __init__(self, members: Sequence[Tuple[Geometry, SpeciesModel, Options]],
                 instances: Optional[Sequence[int]] = None):
        if not members:
            raise ValueError("an ensemble needs at least one member")
        self.setups = [Setup(g, m, o) for g, m, o in members]
        if instances is not None:
            for s, k in zip(self.setups, instances):
                s.s.problem.instance = int(k)
        arr = (CProblem * len(self.setups))(*[s.s.problem for s in self.setups])
        self.handle = C.c_void_p()
        _check(device_lib().cs_batch_create(arr, len(self.setups), C.byref(self.handle)), "cs_batch_create")
        self.count = len(self.setups)
        self.time_ind = 1
        self.ns = int(self.setups[0].s.num_species)
        self.nd = int(self.setups[0].s.num_domains)
        self.mt = self.setups[0].problem.rng == RNG_MT

    @property
    def dt(self):
        return self.setups[0].dt

    def n(self, i: int):
        return self.setups[i].n

    def step(self, nsteps: int = 1):
        lib = device_lib()
        if self.mt:
            # each member's Mersenne Twister draws, gating step by gating
            # step, members one after the other within a gating step
            per = [s.mt_uniforms(self.time_ind, nsteps) for s in self.setups]
            draws = self.draws_per_gate()
            total = sum(draws)
            cnt = C.c_uint64()
            _check(lib.cs_batch_uniforms_needed(self.handle, self.time_ind, nsteps, C.byref(cnt)),
                   "cs_batch_uniforms_needed")
            gates = cnt.value // total if total else 0
            uni = (np.ascontiguousarray(np.hstack([p.reshape(gates, k) for p, k in zip(per, draws)]).ravel())
                   if gates else np.zeros(0))
            _check(lib.cs_batch_step_uniforms(self.handle, self.time_ind, nsteps,
                                              uni.ctypes.data_as(C.POINTER(C.c_double)), len(uni)),
                   "cs_batch_step_uniforms")
        else:
            _check(lib.cs_batch_step(self.handle, self.time_ind, nsteps), "cs_batch_step")
        self.time_ind += nsteps

    def draws_per_gate(self) -> List[int]:
        """Gating uniforms each member draws per gating step (its channels
        without a forced closing time)."""
        out = []
        for s in self.setups:
            p = s.problem
            out.append(sum(int(p.fluxes[f].num_pairs) for f in range(p.num_fluxes)
                           if p.fluxes[f].discrete and not p.fluxes[f].params[8] > 0))
        return out

    def synchronize(self):
        _check(device_lib().cs_batch_synchronize(self.handle), "cs_batch_synchronize")

    def species(self, i: int, s: int) -> np.ndarray:
        out = np.empty(self.n(i), dtype=np.float64)
        _check(device_lib().cs_batch_download_species(self.handle, i, s, out.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_batch_download_species")
        return out

    def set_species(self, i: int, s: int, values: np.ndarray):
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n(i))
        _check(device_lib().cs_batch_upload_species(self.handle, i, s, arr.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_batch_upload_species")

    def stats(self):
        """Per member (min, max, total) arrays of shape [count, nd, ns]."""
        k = self.count * self.nd * self.ns
        mn, mx, tot = np.empty(k), np.empty(k), np.empty(k)
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        _check(device_lib().cs_batch_reduce_stats(self.handle, ptr(mn), ptr(mx), ptr(tot)), "cs_batch_reduce_stats")
        shp = (self.count, self.nd, self.ns)
        return mn.reshape(shp), mx.reshape(shp), tot.reshape(shp)

    def linescans(self, species: int, axis: int, offsets, conv, num_domains: int = 0):
        """Simulation.linescan of every member (one device pass for all)."""
        w, ks = [], []
        for i in range(self.count):
            n = self.n(i)
            h = self.setups[i].problem.h
            off = [(n[ax] // 2 + int(offsets[ax])) * h for ax in range(3)]
            w.extend(gauss(off[ax] - np.arange(n[ax]) * h, conv[ax]) for ax in range(3))
            ks.append(n[axis])
        weights = np.ascontiguousarray(np.concatenate(w))
        proj = np.empty(sum(ks))
        mask = 0xFFFFFFFF if not num_domains else (0xFF if num_domains >= 8 else (1 << num_domains) - 1)
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        _check(device_lib().cs_batch_linescan(self.handle, species, axis, ptr(weights), mask, ptr(proj)),
               "cs_batch_linescan")
        out, o = [], 0
        for i, k in enumerate(ks):
            h = self.setups[i].problem.h
            pos = np.arange(k) * h
            g = gauss(pos[:, None] - pos[None, :], conv[axis])
            out.append(g @ proj[o:o + k])
            o += k
        return out

    def _flux_size(self, i: int, flux: int) -> int:
        return int(self.setups[i].problem.fluxes[flux].num_pairs)

    def channel_states(self, i: int, flux: int) -> np.ndarray:
        n = self._flux_size(i, flux)
        out = np.zeros(max(n, 1), dtype=np.uint8)
        _check(device_lib().cs_batch_get_channel_states(self.handle, i, flux, out.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_batch_get_channel_states")
        return out[:n]

    def recruited(self, i: int, flux: int) -> np.ndarray:
        n = self._flux_size(i, flux)
        out = np.zeros(max(n, 1), dtype=np.uint8)
        _check(device_lib().cs_batch_get_channel_recruited(self.handle, i, flux,
                                                           out.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_batch_get_channel_recruited")
        return out[:n]

    def summary(self):
        """(open channels, time index of the latest gating step with a channel
        open, channels recruited) per member."""
        no = np.zeros(self.count, dtype=np.uint32)
        lo = np.zeros(self.count, dtype=np.uint64)
        nr = np.zeros(self.count, dtype=np.uint32)
        _check(device_lib().cs_batch_channel_summary(self.handle, no.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                     lo.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                     nr.ctypes.data_as(C.POINTER(C.c_uint32))),
               "cs_batch_channel_summary")
        return no, lo, nr

    def set_active(self, active):
        arr = np.ascontiguousarray(np.asarray(active, dtype=bool).astype(np.uint8))
        if arr.size != self.count:
            raise ValueError("one flag per member")
        _check(device_lib().cs_batch_set_active(self.handle, arr.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_batch_set_active")

    def launch(self) -> dict:
        cfg = CBatchLaunch()
        _check(device_lib().cs_batch_get_launch(self.handle, C.byref(cfg)), "cs_batch_get_launch")
        return cfg.as_dict()

    def set_launch(self, **kw):
        cfg = CBatchLaunch(**{**self.launch(), **kw})
        _check(device_lib().cs_batch_set_launch(self.handle, C.byref(cfg)), "cs_batch_set_launch")

    def autotune(self, reps: int = 2):
        cfg, ms = CBatchLaunch(), C.c_double()
        _check(device_lib().cs_batch_autotune(self.handle, self.time_ind, reps, C.byref(cfg), C.byref(ms)),
               "cs_batch_autotune")
        return cfg.as_dict(), ms.value

    def set_graph(self, on: bool):
        _check(device_lib().cs_batch_set_graph(self.handle, 1 if on else 0), "cs_batch_set_graph")

    @property
    def period(self) -> int:
        """Steps of one captured graph period."""
        p = C.c_uint64()
        _check(device_lib().cs_batch_step_period(self.handle, C.byref(p)), "cs_batch_step_period")
        return int(p.value)

    GATE_OWN, GATE_REPLAY, GATE_CLOSURE = 0, 1, 2

    def set_gating(self, mode: int, net: "Optional[Mlp]" = None):
        """Gate as the reference does (GATE_OWN), not at all (GATE_REPLAY:
        states come from copy_channel_states), or on the closure network's
        prediction of each channel mouth's Ca (GATE_CLOSURE)."""
        self._net = net                       # the network must outlive its use
        _check(device_lib().cs_batch_set_gating(self.handle, mode, net.handle if net is not None else None),
               "cs_batch_set_gating")

    def copy_channel_states(self, src: "Ensemble"):
        _check(device_lib().cs_batch_copy_channel_states(self.handle, src.handle), "cs_batch_copy_channel_states")

    @property
    def num_channels(self) -> int:
        n = C.c_uint32()
        _check(device_lib().cs_batch_channel_count(self.handle, C.byref(n)), "cs_batch_channel_count")
        return int(n.value)

    def closure_features(self) -> np.ndarray:
        """Raw closure inputs of every channel, [channels, 30]."""
        out = np.zeros((max(self.num_channels, 1), MLP_IN), dtype=np.float32)
        _check(device_lib().cs_batch_closure_features(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))),
               "cs_batch_closure_features")
        return out[:self.num_channels]

    def channel_ca(self) -> np.ndarray:
        """Absolute Ca at every channel's cleft voxel."""
        out = np.zeros(max(self.num_channels, 1))
        _check(device_lib().cs_batch_channel_ca(self.handle, out.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_batch_channel_ca")
        return out[:self.num_channels]

    def closure_predictions(self) -> np.ndarray:
        out = np.zeros(max(self.num_channels, 1))
        _check(device_lib().cs_batch_closure_predictions(self.handle, out.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_batch_closure_predictions")
        return out[:self.num_channels]

    def channel_sr_ca(self) -> np.ndarray:
        """Absolute Ca at every channel's SR (from) voxel."""
        out = np.zeros(max(self.num_channels, 1))
        _check(device_lib().cs_batch_channel_sr_ca(self.handle, out.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_batch_channel_sr_ca")
        return out[:self.num_channels]

    def channel_open(self) -> np.ndarray:
        """State of every channel (1 open), in channel order."""
        out = np.zeros(max(self.num_channels, 1), dtype=np.uint8)
        _check(device_lib().cs_batch_channel_open(self.handle, out.ctypes.data_as(C.POINTER(C.c_uint8))),
               "cs_batch_channel_open")
        return out[:self.num_channels]

    def set_flux_closure(self, net: "Optional[Mlp]"):
        """With GATE_CLOSURE: scale every channel's flux by the flux
        network's predicted ratio of fine to coarse release (None: off)."""
        self._fnet = net                      # the network must outlive its use
        _check(device_lib().cs_batch_set_flux_closure(self.handle, net.handle if net is not None else None),
               "cs_batch_set_flux_closure")

    def flux_predictions(self) -> np.ndarray:
        out = np.zeros(max(self.num_channels, 1))
        _check(device_lib().cs_batch_flux_predictions(self.handle, out.ctypes.data_as(C.POINTER(C.c_double))),
               "cs_batch_flux_predictions")
        return out[:self.num_channels]

    def monitor(self) -> np.ndarray:
        out = np.zeros(self.count, dtype=np.uint32)
        _check(device_lib().cs_batch_monitor(self.handle, out.ctypes.data_as(C.POINTER(C.c_uint32))),
               "cs_batch_monitor")
        return out

    def channel_owner(self) -> np.ndarray:
        """Realisation of every channel, in channel order."""
        return np.repeat(np.arange(self.count), [sum(self._flux_size(i, f) for f in range(self.setups[i].problem.num_fluxes)
                                                     if self.setups[i].problem.fluxes[f].discrete)
                                                 for i in range(self.count)])

    def time_steps(self, nsteps: int):
        tot, ker = C.c_double(), C.c_double()
        _check(device_lib().cs_batch_time_steps(self.handle, self.time_ind, nsteps, C.byref(tot), C.byref(ker)),
               "cs_batch_time_steps")
        self.time_ind += nsteps
        return tot.value, ker.value

    def info(self):
        cnt, nv, nb = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(device_lib().cs_batch_info(self.handle, C.byref(cnt), C.byref(nv), C.byref(nb)), "cs_batch_info")
        return int(cnt.value), int(nv.value), int(nb.value)

    def reset(self):
        _check(device_lib().cs_batch_reset(self.handle), "cs_batch_reset")
        self.time_ind = 1

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            device_lib().cs_batch_destroy(self.handle)
            self.handle = C.c_void_p()
        for s in getattr(self, "setups", []):
            s.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MLP_NUM_PARAMS = 6337
MLP_IN = 30


def _f32(a, shape=None):
    arr = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        arr = arr.reshape(shape)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The closure's network on the GPU (cs_mlp_*): 30-64-64-1 GELU
    perceptron, bfloat16 matrix-core passes with FP32 accumulation, Adam."""

    def __init__(self, params):
        arr, ptr = _f32(params)
        if arr.size != MLP_NUM_PARAMS:
            raise ValueError(f"expected {MLP_NUM_PARAMS} parameters")
        self.handle = C.c_void_p()
        _check(device_lib().cs_mlp_create(ptr, C.byref(self.handle)), "cs_mlp_create")

    @property
    def params(self) -> np.ndarray:
        out = np.empty(MLP_NUM_PARAMS, dtype=np.float32)
        _check(device_lib().cs_mlp_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))),
               "cs_mlp_get_params")
        return out

    @params.setter
    def params(self, p):
        arr, ptr = _f32(p)
        _check(device_lib().cs_mlp_set_params(self.handle, ptr), "cs_mlp_set_params")

    def set_normalisation(self, in_mean, in_std, in_lo, in_hi, out_mean, out_std, margin):
        a = [_f32(v, (MLP_IN,)) for v in (in_mean, in_std, in_lo, in_hi)]
        _check(device_lib().cs_mlp_set_normalisation(self.handle, a[0][1], a[1][1], a[2][1], a[3][1], float(out_mean),
                                                     float(out_std), float(margin)), "cs_mlp_set_normalisation")

    def forward(self, x) -> np.ndarray:
        xa, xp = _f32(x, (-1, MLP_IN))
        y = np.empty(len(xa), dtype=np.float32)
        _check(device_lib().cs_mlp_forward(self.handle, xp, len(xa), y.ctypes.data_as(C.POINTER(C.c_float))),
               "cs_mlp_forward")
        return y

    def grad(self, x, t):
        xa, xp = _f32(x, (-1, MLP_IN))
        ta, tp = _f32(t, (len(xa),))
        g = np.empty(MLP_NUM_PARAMS, dtype=np.float32)
        loss = C.c_double()
        _check(device_lib().cs_mlp_grad(self.handle, xp, tp, len(xa), g.ctypes.data_as(C.POINTER(C.c_float)),
                                        C.byref(loss)), "cs_mlp_grad")
        return g, loss.value

    def set_data(self, x, t, seed: int = 0):
        xa, xp = _f32(x, (-1, MLP_IN))
        ta, tp = _f32(t, (len(xa),))
        _check(device_lib().cs_mlp_set_data(self.handle, xp, tp, len(xa), int(seed)), "cs_mlp_set_data")

    def train(self, steps: int, batch: int, lr: float) -> float:
        loss = C.c_double()
        _check(device_lib().cs_mlp_train(self.handle, steps, batch, float(lr), C.byref(loss)), "cs_mlp_train")
        return loss.value

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            device_lib().cs_mlp_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calibration_copy(nbytes: int, reps: int) -> float:
    """Plain device copy of nbytes, reps times; returns ms per copy."""
    ms = C.c_double()
    _check(device_lib().cs_calibration_copy(nbytes, reps, C.byref(ms)), "cs_calibration_copy")
    return ms.value


def device_info() -> str:
    dev = C.c_int()
    buf = C.create_string_buffer(256)
    _check(device_lib().cs_device_info(C.byref(dev), buf, 256), "cs_device_info")
    return f"device {dev.value}: {buf.value.decode()}"
