"""ems_amd -- MI355X-native FDTD step engine with the API of the ems engine.

The device engine (HIP, gfx950) sits behind the C ABI of include/ems_gpu.h;
the host side (coefficients, CPML layers, waveforms, run loop) is C++
(include/ems_host.h). This package is the Python mirror of the reference
engine API over both.
"""
from ._native import EmsError, LIB_PATH  # noqa: F401  (fails loudly if unbuilt)
from .engine import (  # noqa: F401
    C0, EPS0, MU0, Z0, Axis, CellMaterials, Cpml, CpmlBoundary, EdgeSource, EndCriteria, Engine,
    EngineBatch, EngineBuilder, Face, FaceCondition, Grid, Material, Probe, ProbeQuantity, RunLimits,
    RunOutcome, Settle, SourceKind, StopReason, TimeStep, Waveform, coefficients,
    courant_limit, cpml_coefficients, drive_script, edge_materials, energy_many, pool_lane)
from .lumped import (  # noqa: F401,E402
    EdgeValues, LumpedBox, LumpedElement, LumpedPort, Topology, fold_parallel)
from .dispersive import Dispersive, Field, Pole, Scheme  # noqa: F401,E402
from .planewave import Medium, PlaneWave, PlaneWaveSpec  # noqa: F401,E402
from .post import PortResult, dft, s_param  # noqa: F401,E402
from .learn import ConvNet, EarlyStop, TraceFormat  # noqa: F401,E402
from .mur import MurBoundary  # noqa: F401,E402
from .upml import OpenEmsPml, openems_pml_sigma  # noqa: F401,E402
from . import openems_probes, raster  # noqa: F401,E402
from .conformal import (  # noqa: F401,E402
    ConformalEdge, ConformalFace, ConformalOptions, ConformalPec, ConformalReport)
