"""bench.py names the kernel it timed (template arguments from the queue and row shape) and looks
its rocprofv3 records up by that name: the names it derives for the stored tuned configurations
must be the kernels the committed profiles traced, or the bench line loses its profile and
traffic figures without an error."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _load(name):
    return json.load(open(os.path.join(ROOT, "profiles", name)))


def test_kernel_names_of_the_tuned_configurations_are_the_profiled_kernels():
    tuned = _load("tuned_configs.json")["1000000x128 nq=10000 k=10 recall>=0.9"]
    profiled = {key.split(" ")[0] for key in _load("kernel_profiles.json")}
    traffic = {key.split(" ")[0] for key in _load("traffic.json")}
    plain = tuned["plain"]
    lanes, chunks = bench._row_shape(128, plain.get("lanes_per_row", 0))
    pname = f"search_kernel<0,{bench._slots(plain['ef'])},{lanes},{chunks},0>"
    lrn = tuned["learned"]
    lanes, chunks = bench._row_shape(128, lrn.get("lanes_per_row", 0))
    mode = 3 if lrn.get("entry_late_hops", 0) > 0 else 1
    lname = f"persistent_kernel<0,{bench._slots(lrn['ef'])},{lanes},{chunks},{mode}>"
    for name in (pname, lname):
        assert name in profiled, (name, profiled)
        assert name in traffic, (name, traffic)


def test_row_shapes_of_the_workloads():
    # the shapes the committed profiles and bench lines show: 128 floats at 16 lanes (2 chunks)
    # or 8 lanes (4 chunks); 960 floats at 64 lanes (4 chunks); 256-bit codes are not float rows
    assert bench._row_shape(128) == (16, 2)
    assert bench._row_shape(128, 8) == (8, 4)
    assert bench._row_shape(960) == (64, 4)
    assert [bench._slots(q) for q in (1, 64, 65, 320, 321, 512)] == [1, 1, 2, 5, 6, 8]
