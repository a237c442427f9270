"""FileReduction (reductions.jl:57-91) on the host: one group trajectory_<id> per trajectory, one dataset per
output, numbers and arrays of numbers as they are, vectors of arrays stacked on a new last axis, in the
layout HDF5.jl gives a Julia array (dimensions reversed, column-major data).  The files are read back with
h5dump, a reader independent of the writer (test/Ensembles/reductions.jl:20-33 checks the same file
shape: OutputPosition three-dimensional, OutputTotalEnergy a vector)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
pytestmark = pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")


def _header(path):
    return subprocess.run([H5DUMP, "-H", path], check=True, capture_output=True, text=True).stdout


def _dataset(path, name, tmp_path, dtype):
    out = tmp_path / (name.replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", str(out), path], check=True, capture_output=True)
    return np.fromfile(out, dtype=dtype)


def _dims(header, group, name):
    block = header.split(f'GROUP "{group}"')[1]
    m = re.search(r'DATASET "%s" \{.*?DATASPACE\s+(SCALAR|SIMPLE \{ \( ([\d, ]+) \))' % re.escape(name), block, re.S)
    assert m, (group, name)
    return () if m.group(1) == "SCALAR" else tuple(int(x) for x in m.group(2).split(","))


def test_filename_extension_rule():
    from nqcd_mi355x import FileReduction
    assert FileReduction("a/b.h5").filename == "a/b.h5"
    assert FileReduction("a/b.hdf5").filename == "a/b.hdf5"
    assert FileReduction("a/b.txt").filename == "a/b.h5"
    assert FileReduction("a/b").filename == "a/b.h5"


def test_layout_and_values(tmp_path):
    from nqcd_mi355x import FileReduction
    rng = np.random.default_rng(3)
    pos = [rng.normal(size=(3, 2)) for _ in range(5)]   # (ndofs, natoms) per save
    pops = [rng.uniform(size=4) for _ in range(5)]
    batch = [{"Time": np.arange(5.0), "OutputPosition": pos, "OutputTotalEnergy": list(rng.normal(size=5)),
              "OutputAdiabaticPopulation": pops, "OutputSurfaceHops": 7, "OutputFinalTime": 12.5,
              "OutputDiscreteState": [1, 2, 2, 1, 1], "Psi": np.array([1 + 2j, 3 - 1j])},
             {"Time": np.arange(5.0), "OutputPosition": [p * 2 for p in pos], "OutputTotalEnergy": [0.0] * 5,
              "OutputAdiabaticPopulation": pops, "OutputSurfaceHops": 0, "OutputFinalTime": 3.0,
              "OutputDiscreteState": [1] * 5, "Psi": np.zeros(2, complex)}]
    red = FileReduction(str(tmp_path / "out.txt"))
    msg, more = red([], batch, [4, 9])
    path = str(tmp_path / "out.h5")
    assert msg == f"Output written to {path}." and more is False
    h = _header(path)
    # Julia (3, 2, 5) -> file (5, 2, 3); vector -> (5,); vector of vectors (4, 5) -> (5, 4); numbers -> scalar
    assert _dims(h, "trajectory_4", "OutputPosition") == (5, 2, 3)
    assert _dims(h, "trajectory_4", "OutputTotalEnergy") == (5,)
    assert _dims(h, "trajectory_4", "OutputAdiabaticPopulation") == (5, 4)
    assert _dims(h, "trajectory_4", "OutputSurfaceHops") == ()
    assert _dims(h, "trajectory_4", "Time") == (5,)
    assert re.search(r'DATASET "OutputSurfaceHops" \{\s+DATATYPE\s+H5T_STD_I64LE', h)
    assert re.search(r'DATASET "Psi" \{\s+DATATYPE\s+H5T_COMPOUND \{\s+H5T_IEEE_F64LE "r";\s+H5T_IEEE_F64LE "i";', h)
    # column-major data: element (i, j, s) of the Julia array at i + 3 j + 6 s
    flat = _dataset(path, "/trajectory_4/OutputPosition", tmp_path, np.float64)
    jl = np.stack(pos, axis=-1)
    assert np.array_equal(flat, jl.flatten(order="F"))
    flat = _dataset(path, "/trajectory_9/OutputPosition", tmp_path, np.float64)
    assert np.array_equal(flat, np.stack([p * 2 for p in pos], axis=-1).flatten(order="F"))
    assert np.array_equal(_dataset(path, "/trajectory_4/OutputAdiabaticPopulation", tmp_path, np.float64),
                          np.stack(pops, axis=-1).flatten(order="F"))
    assert np.array_equal(_dataset(path, "/trajectory_4/OutputDiscreteState", tmp_path, np.int64), [1, 2, 2, 1, 1])
    assert _dataset(path, "/trajectory_4/OutputSurfaceHops", tmp_path, np.int64).tolist() == [7]
    assert _dataset(path, "/trajectory_9/OutputFinalTime", tmp_path, np.float64).tolist() == [3.0]
    txt = subprocess.run([H5DUMP, "-d", "/trajectory_4/Psi", path], check=True, capture_output=True, text=True).stdout
    data = txt.split("DATA {")[1]
    vals = [float(x) for x in re.findall(r"-?\d+(?:\.\d*)?(?:e[-+]?\d+)?", re.sub(r"\(\d+\):", "", data))]
    assert vals == [1, 2, 3, -1]


def test_existing_file_is_appended_and_groups_are_not_overwritten(tmp_path):
    from nqcd_mi355x import FileReduction
    red = FileReduction(str(tmp_path / "f.h5"))
    red([], [{"x": 1.0}], [1])
    red([], [{"x": 2.0}], [2])                       # "cw": the file is opened, trajectory_1 kept
    h = _header(str(tmp_path / "f.h5"))
    assert 'GROUP "trajectory_1"' in h and 'GROUP "trajectory_2"' in h
    with pytest.raises(ValueError, match="already exists"):
        red([], [{"x": 3.0}], [1])


@pytest.mark.parametrize("value", [{"reflection": np.zeros(2), "transmission": np.ones(2)}, "text",
                                   [np.zeros(2), np.zeros(3)]])
def test_unconvertible_outputs_raise(tmp_path, value):
    from nqcd_mi355x import FileReduction
    with pytest.raises(ValueError, match="Cannot convert output type to HDF5 format"):
        FileReduction(str(tmp_path / "g.h5"))([], [{"y": value}], [1])
