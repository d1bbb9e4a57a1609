"""qsl's on-disk results format (Results<T,R>::writeToFile,
include/qsl/benchmark/results.tpp:316-355), pinned by a file the reference
itself wrote: tests/golden/np_phase.csv is the reference's
results/np_phase.csv (examples/benchmark-np.cpp:9-88).  Both writers here,
the Python one (qsl_amd/results.py) and the C++ one
(include/qsl/benchmark/hip_results.hpp, through tests/cpp/results_echo),
must reproduce it byte for byte from its parsed values.  On an MI355X the
sweep harness (include/qsl/benchmark/hip_sweep.hpp, tests/cpp/sweep_hip)
writes a file of the same format."""
import os
import subprocess

import numpy as np
import pytest

from qsl_amd import results as R

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "np_phase.csv")
ECHO = os.path.join(HERE, "cpp", "results_echo")
SWEEP = os.path.join(HERE, "cpp", "sweep_hip")


def _golden_text():
    with open(GOLDEN) as f:
        return f.read()


def test_python_writer_reproduces_reference_file():
    text = _golden_text()
    r = R.read(GOLDEN)
    assert r.getMeta() == ["Testing the phase gate"]
    assert r.headings() == ["nones", "Default", "Default+OMP", "NP", "NP+OMP"]
    assert r.column(0) == list(range(1, 16))
    assert r.text() == text
    # the statistics of Results (results.tpp:139-219)
    col = np.array(r.column("Default"))
    assert abs(r.mean("Default") - col.mean()) < 1e-18
    assert abs(r.variance(1) - col.var(ddof=1)) < 1e-20


def test_python_writer_refuses_bad_rows(tmp_path):
    r = R.Results(["n", "a", "b"])
    with pytest.raises(ValueError):
        r.addRow([1], [0.5])  # 2 columns under 3 headings
    r.addRow([1], [0.5, 0.25])
    with pytest.raises(ValueError):
        r.addRow([1, 2], [0.5])
    with pytest.raises(ValueError):
        r.addRow([1], [0.5, 0.25, 1.0])
    r.addMeta("meta")
    p = tmp_path / "t.csv"
    r.writeToFile(str(p))
    assert p.read_text() == "# meta\n\nn a b\n1 0.5 0.25\n"
    assert R.read(str(p)).text() == p.read_text()


def test_cpp_writer_reproduces_reference_file():
    if not os.path.exists(ECHO):
        pytest.fail(f"{ECHO} not built: run __graft_entry__.build()")
    r = R.read(GOLDEN)
    lines = [str(len(r.meta))] + r.meta + [" ".join(r.head),
                                            "%d %d" % (1, len(r.intcols))]
    for ints, fps in zip(r.intcols, r.fpcols):
        lines.append(" ".join([str(v) for v in ints] + [repr(v) for v in fps]))
    out = subprocess.run([ECHO], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, timeout=30)
    assert out.returncode == 0, out.stderr
    assert out.stdout == _golden_text()


@pytest.mark.gpu
def test_sweep_writes_results_file(tmp_path):
    if not os.path.exists(SWEEP):
        pytest.fail(f"{SWEEP} not built: run __graft_entry__.build()")
    for gate, nmin in (("rotateX", 2), ("controlNot", 2)):
        out = tmp_path / f"sweep_{gate}.csv"
        r = subprocess.run([SWEEP, gate, str(out), "12", str(nmin), "14"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        res = R.read(str(out))
        assert res.headings() == ["nqubits", "Qub<hip,double>", "Qub<hip,float>"]
        assert res.getMeta() == [f"Averaged gate times for {gate}", "Results are in seconds"]
        assert res.column(0) == list(range(nmin, 14))
        for c in (1, 2):
            t = np.array(res.column(c))
            assert np.all(t > 0) and np.all(t < 1e-2), t
        # the file is exactly what the writer makes of its own values
        assert res.text() == out.read_text()
