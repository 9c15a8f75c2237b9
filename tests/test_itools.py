"""Interaction-dump format (exadem_amd.itools) against a dump the reference wrote.

tests/golden/interaction_dump_8000_0.txt is the reference's own test data file
scripts/test/data/compute_clusters_stresses_tensors/InteractionOutputDir-8000/InteractionOutputDir-8000_0.txt
(78 polyhedron interactions written by itools::create_buffer / write_file, writer.hpp:35-115). Parsing it
and writing the values back must reproduce every line character for character: column order, sub-feature
and type columns, and the default C++ ostream number format.
"""
import os

import numpy as np

from exadem_amd import itools

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "interaction_dump_8000_0.txt")


def test_reference_dump_round_trip():
    keys, vals = itools.read_dump(GOLDEN)
    assert keys.shape == (78, 5) and vals.shape == (78, 16)
    lines = itools.format_rows(keys[:, 0], keys[:, 1], keys[:, 2], keys[:, 3], keys[:, 4], vals)
    with open(GOLDEN) as f:
        ref = [line.rstrip("\n") for line in f if line.strip()]
    assert lines == ref


def test_number_format_matches_cpp_default():
    # std::ostream << double with the default precision 6 and no floatfield
    cases = [(0.0, "0"), (-0.0, "-0"), (1.0, "1"), (0.1, "0.1"), (123456.0, "123456"), (1234567.0, "1.23457e+06"),
             (-9.47075e-07, "-9.47075e-07"), (1e-5, "1e-05"), (0.0001, "0.0001"), (2.5e100, "2.5e+100")]
    for x, s in cases:
        assert itools.format_rows([1], [2], [0], [0], [0], np.full((1, 16), x))[0].split(",")[5] == s, x


def test_dump_path_layout(tmp_path):
    p = itools.dump_path(str(tmp_path), "InteractionOutputDir-", 8000, rank=0)
    assert p == os.path.join(str(tmp_path), "ExaDEMAnalyses", "InteractionOutputDir-8000",
                             "InteractionOutputDir-8000_0.txt")
