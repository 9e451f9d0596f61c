"""The package exposes the engine API the tests and the benchmark use
(the names of the reference's engine module, src/engine/mod.rs)."""
import ems_amd as E


def test_public_names():
    for name in ("Engine", "EngineBatch", "EngineBuilder", "Grid", "CellMaterials", "Material",
                 "CpmlBoundary", "Cpml", "EdgeSource", "Probe", "Waveform", "RunLimits",
                 "EndCriteria", "LumpedBox", "LumpedElement", "LumpedPort", "Topology",
                 "EdgeValues", "fold_parallel", "Dispersive", "Pole", "Field", "Scheme",
                 "PlaneWave", "PlaneWaveSpec", "Medium", "PortResult", "dft", "s_param",
                 "ConformalPec", "ConformalEdge", "ConformalFace", "ConformalOptions",
                 "ConvNet", "EarlyStop", "TraceFormat"):
        assert hasattr(E, name), name
