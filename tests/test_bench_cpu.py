"""bench.py's contract on the CPU: one GPU only (a clear message and exit code 2 otherwise, before
anything touches a GPU) and the workloads its configurations name."""
import json

import numpy as np

import bench


def test_more_than_one_gpu_is_refused_with_a_message(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    line = json.loads(capsys.readouterr().out.strip())
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_configurations_are_the_named_workloads():
    c3 = bench.make_params(bench.CONFIGS["c3"])
    assert c3["NumberOfImage"] == [512, 512, 512] and c3["ImagePixelSpacing"] == [0.25] * 3
    assert c3["NumberOfViews"] == 720 and c3["NumberOfDetectorPixels"] == [1024, 768]
    assert c3["DetectorPixelSize"] == [0.4, 0.4] and c3["FilterType"] == "shepp-logan"
    assert c3["SourceInit"] == [0, 1000.0, 0] and c3["DetectorInit"] == [0, -500.0, 0]
    assert c3["EndAngle"] == 2 * np.pi
    c2 = bench.make_params(bench.CONFIGS["c2"])
    assert c2["NumberOfImage"] == [256, 256, 256] and c2["NumberOfViews"] == 360
    assert c2["NumberOfDetectorPixels"] == [384, 384]
    c1 = bench.make_params(bench.CONFIGS["c1"])
    assert c1["NumberOfImage"] == [256, 256, 1] and c1["NumberOfDetectorPixels"] == [512, 1]
    assert c1["DetectorShape"] == "Curved" and c1["FilterType"] == "ram-lak"


def test_pixel_leg_runs_the_reference_geometry_file():
    """The pixel-driven leg's geometry is python/ReconstructionParams.txt (the committed fixture)
    as the params-file reader returns it, data path aside."""
    import os

    from ctreconstruction_amd import paramsfile

    golden = os.path.join(os.path.dirname(__file__), "golden", "ReconstructionParams.txt")
    read = paramsfile.read_params_file(golden)
    for key, value in bench.PIXEL_GEOMETRY.items():
        if key != "DataPath":
            assert read[key] == value, key
