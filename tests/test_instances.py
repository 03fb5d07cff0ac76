"""All 30 reference instances: fixture loading, and the reader of the
reference's own data/ tree against the fixtures written from it."""
import json
import os

import numpy as np
import pytest

from brkga_fft2_amd.instances import machines_of, problem_from_data
from brkga_fft2_amd.problem import pack_problem, thresholds_for
from tests.helpers import GOLDEN, ref_instance

with open(os.path.join(GOLDEN, "ref_instances.json")) as _f:
    NAMES = list(json.load(_f))

REF_DATA = "/root/reference/data"


def test_fixture_lists_every_instance():
    assert len(NAMES) == 30
    sizes = sorted({(int(n[1:].split("M")[0]), machines_of(n)) for n in NAMES})
    assert sizes == [(25, 2), (50, 2), (75, 2), (100, 4), (150, 4), (200, 4)]


@pytest.mark.parametrize("name", NAMES)
def test_instance_loads(name):
    prob, n, M, packed = ref_instance(name)
    assert n == int(name[1:].split("M")[0]) and M == machines_of(name)
    assert packed["n_parts"] == n and packed["n_machines"] == M
    assert len(packed["slot_type"]) == n
    # every rotation of every part type fits its mask block
    assert int(packed["mask_off"][-1]) + int(packed["rot_h"][-1]) * int(packed["rot_w"][-1]) == len(packed["masks"])


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data/ tree is not here")
@pytest.mark.parametrize("name", ["P25M2-0", "P100M4-2", "P200M4-4"])
def test_data_reader_matches_fixtures(name):
    prob_f, n, M, packed_f = ref_instance(name)
    prob_d, n_d, M_d = problem_from_data(REF_DATA, name)
    packed_d = pack_problem(prob_d, n_d, M_d, thresholds_for(M_d), prob_d.instance_parts)
    assert (n_d, M_d) == (n, M)
    for k, v in packed_f.items():
        np.testing.assert_array_equal(np.asarray(packed_d[k]), np.asarray(v), err_msg=k)
