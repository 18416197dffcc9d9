"""The benchmark workloads' generated inputs are the committed ones.

bench.py's generators (numpy default_rng streams, the FCC lattice, the nearest-neighbour segments, the
icosphere, the trajectory, the request stream) must keep producing the arrays whose SHA-256 digests
tests/golden/make_workload_inputs.py recorded in tests/golden/workload_inputs.json; the bench line reports
the same digests for the workload it measured (config.inputs_sha256)."""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

import make_workload_inputs  # noqa: E402


def test_generated_inputs_match_committed_digests():
    with open(os.path.join(HERE, 'golden', 'workload_inputs.json')) as f:
        expected = json.load(f)
    assert make_workload_inputs.digests() == expected
