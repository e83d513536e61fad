"""Golden fixtures (tests/golden/oracle_fixtures.npz, written by
tests/golden/golden.py from the FP64 oracle after its closed-form checks).

CPU: the oracle re-run reproduces every fixture (stored as float32, so rel
L2 <= 1e-6): the oracle has not drifted.  GPU: the HIP path matches the
fixtures without running the oracle: 3D model-pass records and fields (and
every field's norm) rel L2 <= 1e-5, 2D records / fields / Born data
<= 1e-5 and the adjoint (a sum over all steps of products of two FP32
wavefields) <= 1e-4, as in the live parity tests.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import golden  # noqa: E402

import cases  # noqa: E402


@pytest.fixture(scope="module")
def fixtures():
    with np.load(golden.FIXTURES) as d:
        return {k: d[k] for k in d.files}


def test_oracle_reproduces_fixtures(fixtures):
    out = golden.oracle_outputs()
    assert sorted(out) == sorted(fixtures)
    for k, ref in fixtures.items():
        e = cases.rel_l2(out[k], ref)
        assert e <= 1e-6, (k, e)


class _GPU3D:
    def __init__(self, p):
        import importlib
        erm = importlib.import_module("2d_and_3d_elastic_reverse_time_migration_amd")
        self.g = erm.RTM3D(p)

    def load_model(self, which, *a):
        self.g.load_model(which, *a)
        if which == 0:
            self.g.zero_wavefield()

    def step(self, *a):
        self.g.step(*a)

    def get_field(self, k):
        return self.g.get_field(k)

    def get_record(self):
        return self.g.get_record()


class _GPU2D:
    def __init__(self, p):
        import importlib
        erm = importlib.import_module("2d_and_3d_elastic_reverse_time_migration_amd")
        self.g = erm.Elastic2D(p)

    def __getattr__(self, k):
        return getattr(self.g, k)

    def forward_record(self):
        return self.g.forward()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["d4", "c4", "c6"])
def test_gpu3d_matches_fixtures(fixtures, name):
    print(f"golden 3D case {name}", flush=True)
    out = golden.outputs_3d(name, _GPU3D)
    for k, a in out.items():
        e = cases.rel_l2(a, fixtures[k])
        assert e <= 1e-5, (k, e)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["e2", "v2"])
def test_gpu2d_matches_fixtures(fixtures, name):
    print(f"golden 2D case {name}", flush=True)
    out = golden.outputs_2d(name, _GPU2D)
    for k, a in out.items():
        e = cases.rel_l2(a, fixtures[k])
        assert e <= (1e-4 if k.endswith("adjoint") else 1e-5), (k, e)
