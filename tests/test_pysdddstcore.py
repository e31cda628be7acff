"""The PySdddstCore surface (reference src/bindings.cpp:49-210) on the CPU:
module functions, PrecisionHandler against the golden sequences made from
the reference's own precision_handler.cpp, SimulationData's properties,
vectors and files, the field / protocol objects and their release into a
SimulationData, and the stress protocols against the golden table.  Nothing
here touches a device (the Simulation itself is exercised by
test_gpu_pysdddstcore.py)."""
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from sdddst_amd import PySdddstCore as sd
from sdddst_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_module_functions():
    assert sd.hello_world() == "Hello World!"
    lib = O.oracle()
    for v in (-3.7, -0.5, -0.25, 0.0, 0.4999999, 0.5, 1.25, 17.5):
        assert sd.normalize(v) == lib.orc_normalize(v)
    assert sd.normalize("text") == "text"
    for x in (-0.3, 0.01, 0.2, 0.49):
        assert sd.X(x) == math.sin(2.0 * math.pi * x) * 0.5 / math.pi
        assert sd.X2(x) == (1.0 - math.cos(2.0 * math.pi * x)) * 0.5 / math.pi / math.pi
        assert sd.X_dx(x) == math.cos(2.0 * math.pi * x)
        assert sd.X2_dx(x) == math.sin(2.0 * math.pi * x) / math.pi
        K = 1.65 * 1.65 * 1e6 / 256.0
        assert sd.E(x, 0.1, K) == math.exp(-K * (sd.X2(x) + sd.X2(0.1)))
        assert sd.E_dx(x, 0.1, K) == -sd.E(x, 0.1, K) * K * sd.X2_dx(x)
    assert sd.get_wall_time() > 1e9 and sd.get_cpu_time() >= 0.0


def test_precision_handler_golden_sequences():
    g = np.load(os.path.join(GOLDEN, "precision_handler.npz"))
    ro = eo = 0
    for k in range(len(g["size"])):
        size, prec = int(g["size"][k]), float(g["prec"][k])
        r2 = g["r2"][ro:ro + g["r2_len"][k]]
        ids = g["ids"][ro:ro + g["r2_len"][k]]
        ro += g["r2_len"][k]
        err = g["err"][eo:eo + g["err_len"][k]]
        eo += g["err_len"][k]
        ph = sd.PrecisionHandler()
        ph.set_min_precisity(prec)
        ph.set_size(size)
        assert ph.get_size() == size and ph.get_min_precisity() == prec
        for r, i in zip(r2, ids):
            ph.update_tolerance(float(r), int(i))
        for i, e in enumerate(err):
            ph.update_error(abs(float(e)), i)
        assert ph.get_max_error_ratio_square() == g["max_ratio"][k]
        assert ph.get_new_stepsize(float(g["old"][k])) == g["new_step"][k]
        ph.reset()
        assert ph.get_max_error_ratio_square() == 0.0


def _write(tmp_path, n=40, pc=5, seed=3):
    x, y, b = configs.dislocations(n, seed)
    px, py = configs.point_defects(n, pc, seed)
    d, p = tmp_path / "in.dconf", tmp_path / "pd.dconf"
    configs.write_dconf(str(d), x, y, b)
    with open(p, "w") as f:
        for u, v in zip(px, py):
            f.write("%.16e %.16e\n" % (u, v))
    return d, p, configs.read_dconf(str(d))


def test_simulation_data_properties_and_files(tmp_path):
    d, p, (x, y, b) = _write(tmp_path)
    obj = sd.SimulationDataObject(str(d), str(p))
    data = obj()
    assert data is obj() and str(obj) == "Simulation Data Container Object"
    assert data.dislocation_count == 40 and data.point_defect_count == 5
    dis = data.dislocations
    assert isinstance(dis, sd.DislocationVector) and len(dis) == 40
    assert dis[3] == sd.Dislocation(x[3], y[3], b[3])
    assert str(dis[0]) == "x: %f y: %f b: %f" % (x[0], y[0], b[0])
    assert len(data.point_defects) == 5
    # defaults of SimulationData (simulation_data.cpp:43-117)
    assert data.iteration_count == 2 and data.simulation_time == 0.0 and data.successful_steps == 0
    assert data.step_count_limited is False and data.time_limited is False
    # the cutoff rule (simulation_data.cpp:208-218)
    data.cutoff_multiplier = 0.5
    data.update_cutoff()
    rc, rc2, inv = O.cutoff(40, 0.5)
    assert data.cutoff == rc and data.cutoff_square == rc2 and data.one_per_cutoff_square == inv
    # writes read back
    data.step_size = 3e-7
    data.step_count_limit = 17
    data.step_count_limited = True
    data.avalanche_speed_threshold = 0.25
    data.final_dislocation_configuration_path = str(tmp_path / "out.dconf")
    assert data.step_size == 3e-7 and data.step_count_limit == 17 and data.step_count_limited is True
    assert data.avalanche_speed_threshold == 0.25
    assert data.final_dislocation_configuration_path == str(tmp_path / "out.dconf")
    with pytest.raises(Exception):
        data.dislocation_count = 3
    with pytest.raises(AttributeError):
        data.no_such_property = 1
    # dislocations replaced and written out
    data.dislocations = [sd.Dislocation(0.1, 0.2, 1.0), sd.Dislocation(-0.3, 0.4, -1.0)]
    assert data.dislocation_count == 2
    data.write_dislocation_data_to_file(str(tmp_path / "two.dconf"))
    xs, ys, bs = configs.read_dconf(str(tmp_path / "two.dconf"))
    assert xs.tolist() == [0.1, -0.3] and bs.tolist() == [1.0, -1.0]
    data.read_dislocation_data_from_file(str(d))
    assert data.dislocation_count == 40
    data.write_point_defect_data_to_file(str(tmp_path / "pd2.dconf"))
    assert np.array_equal(np.loadtxt(str(tmp_path / "pd2.dconf")), np.loadtxt(str(p)))
    # the per-dislocation vectors are sized with the dislocations
    assert len(data.init_speed) == 40 and len(data.d_vec) == 40
    data.g_vec = list(range(40))
    assert data.g_vec[39] == 39.0


def test_malformed_configuration_is_an_error(tmp_path):
    bad = tmp_path / "bad.dconf"
    bad.write_text("0.1 0.2 1\n0.3 0.4\n")
    with pytest.raises(Exception, match="whole records"):
        sd.SimulationData(str(bad), "")
    bad.write_text("0.1 0.2 x\n")
    with pytest.raises(Exception, match="not a number"):
        sd.SimulationData(str(bad), "")


def test_field_and_protocol_objects_are_released_into_the_data(tmp_path):
    d, p, _ = _write(tmp_path)
    data = sd.SimulationData(str(d), "")
    fo = sd.AnalyticFieldObject()
    assert not fo.valid() and fo.name() == "Analytic Field, number of images: 4"
    fo.init()
    assert fo.valid()
    data.tau = fo
    assert not fo.valid()                      # released (simulation_data.cpp:227-235)
    proto = sd.FixedRateStressProtocolObject()
    proto.rate = 5.0
    assert proto.rate == 0.0                   # not initialised: no protocol to set
    proto.init()
    proto.rate = 5.0
    assert proto.rate == 5.0 and str(proto) == "Fixed Rate Protocol - valid: true"
    data.external_stress = proto
    assert not proto.valid()
    es = data.external_stress
    es.calculate_stress(0.25, data.dislocations, sd.StepType.EndOfBigStep)
    assert es.get_stress(sd.StepType.EndOfBigStep) == 0.25 * 5.0
    assert es.get_stress(sd.StepType.Original) == 0.0


def test_stress_protocols_golden_table(tmp_path):
    # the table made from the reference's own protocol translation units
    # (kind 0 zero, 1 fixed rate, 2 spring); the spring reads the strain the
    # SimulationData accumulated
    g = np.load(os.path.join(GOLDEN, "stress_protocols.npz"))["table"]
    d, p, _ = _write(tmp_path)
    data = sd.SimulationData(str(d), "")
    lib = sd._lib()
    for kind, rate, spring, t, strain, typ, val in g:
        assert lib.sdddst_data_set_protocol(data._h, int(kind), float(rate), float(spring)) == 0
        data.total_accumulated_strain_increase = float(strain)
        if kind == 2:
            # SpringProtocol::calculateStress takes the strain as an argument
            # in the simulator; the bindings' calculate_stress has none, so the
            # spring rows are checked through the simulator's own path (GPU
            # run-loop tests) and only the others here
            continue
        es = data.external_stress
        es.calculate_stress(float(t), data.dislocations, sd.StepType(int(typ)))
        assert es.get_stress(sd.StepType(int(typ))) == val
