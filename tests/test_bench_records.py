"""Benchmark records in the reference's schema (benchmarks/bench_common.h:42-87).

Fixtures: the reference's own result records of benchmarks/bench_nqs.c
(benchmarks/results/nqs/*.json, copied as data into
tests/golden/bench_nqs_reference/).  The records this library writes must
carry the same top-level fields in the same order, the same suite and name,
the same metric keys in the same order and the same workload sizes; the
rate metric is this library's own measurement.
"""
import glob
import json
import os

import pytest

from spin_based_neural_network_amd import bench_records

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bench_nqs_reference")
TOP = ["suite", "name", "utc_epoch", "os", "arch", "hostname", "metrics"]
RATES = {"samples_per_second", "sr_steps_per_second", "local_energy_per_second"}


def _check(rec, ref):
    assert list(rec) == TOP == list(ref)
    assert rec["suite"] == ref["suite"] and rec["name"] == ref["name"]
    assert isinstance(rec["utc_epoch"], int)
    assert list(rec["metrics"]) == list(ref["metrics"])
    for k, v in ref["metrics"].items():
        if k in RATES:
            assert rec["metrics"][k] > 0
        else:
            assert rec["metrics"][k] == v and isinstance(rec["metrics"][k], int), k


def test_emitter_matches_reference_schema(tmp_path):
    """The emitter alone (no device): a record with the sampler_L4 fields."""
    ref = json.load(open(os.path.join(REF, "sampler_L4.json")))
    path = bench_records._emit(str(tmp_path), "nqs", "sampler_L4",
                               [("L", 4), ("samples", 4096), ("samples_per_second", 1.25e6)])
    assert path == os.path.join(str(tmp_path), "nqs", "sampler_L4.json")
    _check(json.load(open(path)), ref)
    assert '"samples_per_second": 1250000\n' in open(path).read()   # %.9g


@pytest.mark.gpu
def test_records_of_the_reference_workloads(tmp_path, hip):
    """The bench_nqs.c workloads through libnqs_hip.so: every reference
    record has a counterpart with the same schema and sizes."""
    paths = bench_records.write_records(str(tmp_path), lib=hip)
    names = {os.path.basename(p)[:-5] for p in paths}
    refs = sorted(glob.glob(os.path.join(REF, "*.json")))
    assert len(refs) == 6
    for r in refs:
        name = os.path.basename(r)[:-5]
        assert name in names, name
        _check(json.load(open(os.path.join(str(tmp_path), "nqs", name + ".json"))), json.load(open(r)))
