"""Write tests/golden/workload_inputs.json: the SHA-256 of every input array the benchmark workloads
generate (bench.py's generators: C1 4096 spheres, C2 the 1,048,576-sphere lattice, C3 the polyhedron /
spherocylinder / mesh mix, the first and last C4 trajectory frames, and the C5 request arrays), so a
change to a generator, or to numpy's generators, shows up as a failing test instead of a silently
different benchmark.

    python tests/golden/make_workload_inputs.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import bench  # noqa: E402


def digests():
    out = {'c1': bench.input_digests(bench.c1_arrays()), 'c2': bench.input_digests(bench.c2_arrays()),
           'c3': bench.input_digests(bench.c3_arrays())}
    frames = bench.trajectory_frames(64)
    out['c4'] = bench.input_digests({'frame_0': frames[0], 'frame_63': frames[-1]})
    req = {}
    for i, spec in enumerate(bench.make_requests(24)):
        for k, v in bench.request_arrays(spec).items():
            req[f'request_{i:02d}_{k}'] = v
    out['c5'] = bench.input_digests(req)
    return out


if __name__ == '__main__':
    with open(os.path.join(HERE, 'workload_inputs.json'), 'w') as f:
        json.dump(digests(), f, indent=1, sort_keys=True)
        f.write('\n')
