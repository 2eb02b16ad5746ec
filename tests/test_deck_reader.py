"""The library's C++ .inp reader (svx_deck_parse, used by the request
server's host workers) against the Python reader solvix_amd/inp.py on the
same text: the golden decks, decks written from generated models, and a
hand-written deck that exercises every rule of the subset (parser.rs:87-1371:
sets and GENERATE, element-set merging, continuation lines, steps with
OP=NEW, dof ranges, loads on node sets, case, comments, CRLF, skipped
keywords and element types). Both readers must give the same descriptor
arrays, and the same line and message on errors. No device is needed."""
import os

import numpy as np
import pytest

from solvix_amd import decks
from solvix_amd.core import parse_deck
from solvix_amd.inp import ParseError, parse_inp, write_inp
from solvix_amd.model import SolveError

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _arrays(model):
    """Name-free view of a model: the descriptor the C ABI sees."""
    d, keep = model.to_desc()
    out = {}
    for k in ("node_ids", "coords", "elem_ids", "kinds", "conn_ptr", "conn_ids", "cload_node", "cload_dof",
              "cload_value"):
        out[k] = np.asarray(getattr(model, k))
    out["youngs"] = np.array([m[1] for m in model.materials])
    out["poisson"] = np.array([m[2] for m in model.materials])
    out["elsets"] = [np.asarray(s[1]).tolist() for s in model.elsets]
    names = [s[0] for s in model.elsets]
    mats = [m[0] for m in model.materials]
    out["assign"] = [(names.index(a) if a in names else -1, mats.index(b) if b in mats else -1)
                     for a, b in model.assignments]
    out["thickness"] = model.thickness
    out["bcs"] = [(np.asarray(n).tolist(), np.asarray(dd).tolist()) for n, dd in model.bcs]
    return out


def _same(text):
    a, b = _arrays(parse_inp(text)), _arrays(parse_deck(text))
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k
    return a


@pytest.mark.parametrize("name", ["cantilever", "cantilever3d", "plate_hole"])
def test_golden_decks(name):
    a = _same(open(os.path.join(GOLDEN, f"{name}.inp")).read())
    assert a["node_ids"].size > 0 and a["elem_ids"].size > 0


def test_written_decks_of_generated_models():
    models = [decks.cfg1(nx=40, ny=4), decks.cfg2(nx=8, ny=2, nz=2), decks.readme_cantilever(20, 4),
              decks.solid_instance(decks.SplitMix64(5), 3e3, 5e3)]
    models += [decks.plane_instance(decks.SplitMix64(100 + s), 1e3, 2e3) for s in range(6)]  # CPS3 and CPS4
    for m in models:
        a = _same(write_inp(m))
        assert a["elem_ids"].size == m.n_elem and np.array_equal(a["coords"], m.coords)


RULES = """** every rule of the subset, mixed case
*Heading
  rules deck
*NODE
1, 0.0, 0.0
2, 1.0, 0.0, 0.0
3, 2.0, 0.0
4, 0.0, 1.0
5, 1.0, 1.0
6, 2.0, 1.0
7, 0.5, 2.0
*element, type=CPS4, elset=Plate
1, 1, 2, 5, 4
2, 2, 3,
   6, 5
*ELEMENT, TYPE=CPS3, ELSET=top
3, 4, 5, 7
*ELEMENT, TYPE=B31, ELSET=BEAMS
9, 1, 2
*Elset, elset=PLATE
3
*ELSET, ELSET=ALLE, GENERATE
1, 3, 2
*ELSET, ELSET=ALLE
2, 2
*NSET, NSET=left
1, 4
*NSET, NSET=TIP, GENERATE
3, 6, 3
*ORIENTATION, NAME=IGNORED
1., 0., 0., 0., 1., 0.
*MATERIAL, NAME=steel
*DENSITY
7850.
*ELASTIC
210e9, 0.3
*MATERIAL, NAME=Steel
*ELASTIC
1.0, 0.1
*MATERIAL, NAME=alu
*ELASTIC
70e9, 0.33
*SOLID SECTION, ELSET=missing, MATERIAL=STEEL
0.5
*SOLID SECTION, ELSET=top, MATERIAL=ALU
0.02
*SOLID SECTION, ELSET=plate, MATERIAL=steel
0.01
*SOLID SECTION, ELSET=ALLE, MATERIAL=nothing
*BOUNDARY
LEFT, 1, 2
*CLOAD
TIP, 2, -5.0
*STEP
*STATIC
*CLOAD
7, 1, 3.5
*END STEP
*STEP, NAME=second
*BOUNDARY
3, 2
*CLOAD, OP=NEW
tip, 1, 2.0e3
6, 2, -1.5
*END STEP
"""


def test_every_rule_of_the_subset():
    a = _same(RULES)
    assert a["elem_ids"].tolist() == [1, 2, 3]  # the B31 element is skipped
    assert a["conn_ids"].tolist() == [1, 2, 5, 4, 2, 3, 6, 5, 4, 5, 7]  # continuation line
    assert a["elsets"] == [[1, 2, 3], [3], [1, 3, 2, 2]]  # PLATE merged; GENERATE with step; duplicates kept
    assert a["assign"] == [(-1, 0), (1, 2), (0, 0), (2, -1)]  # first material of a name; missing ones
    assert a["thickness"] == 0.01  # the last section with data
    assert a["bcs"] == [([1, 4], [1, 2]), ([3], [2])]  # step 2 keeps step 1's (and the model's) constraints
    assert a["cload_node"].tolist() == [3, 6, 6]  # OP=NEW dropped the earlier loads
    assert a["cload_value"].tolist() == [2000.0, 2000.0, -1.5]


def test_crlf_and_blank_lines():
    text = open(os.path.join(GOLDEN, "plate_hole.inp")).read()
    crlf = "\r\n\r\n".join(text.splitlines()) + "\r\n"
    a, b = _arrays(parse_deck(text)), _arrays(parse_deck(crlf))
    for k in a:
        assert (np.array_equal(a[k], b[k]) if isinstance(a[k], np.ndarray) else a[k] == b[k]), k
    _same(crlf)


@pytest.mark.parametrize("bad", [
    "*NODE\n1, 0.0\n",
    "*NODE\n1, 0, 0\n2, 1, 0\n*ELEMENT, TYPE=CPS4\n1, 1, 2, 3\n",
    "*NODE\n1, 0, 0\n*CLOAD\n1, 2\n",
    "*NODE\n1, 0, 0\n*BOUNDARY\nNOSUCHSET, 1\n",
    "*NODE\n1, 0, 0\n*NSET, NSET=A\n1\n*CLOAD\nb, 1, 1.0\n",
])
def test_errors_give_the_same_line_and_message(bad):
    with pytest.raises(ParseError) as py:
        parse_inp(bad)
    with pytest.raises(SolveError) as cpp:
        parse_deck(bad)
    assert cpp.value.kind == "Parse"
    assert cpp.value.message == str(py.value)
