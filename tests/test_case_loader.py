"""Case-directory parsing (fields.json + <field>.json), after tests/test_boundary.cpp:151-212
on the reference's own fixtures (tests/cases/boundary)."""
import os

import pytest

from prism_amd import case

BOUNDARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "boundary")


def read(name):
    return case.read_boundary(os.path.join(BOUNDARY, name, "fields.json"))


def test_good_one_field():
    fields, bcs = read("good_one_field")
    assert [(f[0], f[1]) for f in fields] == [("T", "scalar")]
    assert len(bcs) == 3


def test_good_two_fields():
    fields, bcs = read("good_two_fields")
    assert [f[0] for f in fields] == ["T", "U"]
    assert fields[0][2] == "greenGauss" and fields[0][3] is not None and len(fields[0][3]) == 9
    assert fields[1][2] == "leastSquares"
    assert len(bcs) == 3
    assert bcs["Inlet"]["T"] == ("fixed", 350.0)
    assert bcs["Inlet"]["U"] == ("fixed", (1.0, 0.0, 0.0))
    assert bcs["Wall"]["U"] == "noSlip"


def test_good_empty_fields():
    fields, bcs = read("good_empty_fields")
    assert fields == [] and bcs == {}


@pytest.mark.parametrize("name", ["bad_missing_name", "bad_missing_type", "bad_missing_field_file",
                                  "bad_inconsistent_patches"])
def test_bad_cases_raise(name):
    with pytest.raises(case.CaseError):
        read(name)
