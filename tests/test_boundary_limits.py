"""What the stateless pair driver accepts and rejects, checked before any device work
(N = 0 or a rejected argument: no GPU needed).

* Per-type-pair tables that differ in any field are accepted (the reference's evaluators
  take a param_type per (ti, tj), oxdna/forces.py len_keys=2).
* ntypes is limited to 1..4: the oxDNA type set A, G, C, T/U (reference oxdna/io.py:20),
  which the bonded stacking table st_eps[4][4] also assumes
  (src/OxDNABondedParams.h:198-201); more types are rejected with that reason.
* Negative or NaN squared cut-offs and unknown term ids are rejected.
"""

import ctypes as C

import numpy as np
import pytest

from oxdna import _abi, _engine
from oxdna.model import dna2


@pytest.fixture(scope="module")
def lib():
    return _engine.load()


def _call(lib, term, ntypes, table, rcutsq, n=0):
    box = _abi.Box()
    box.L[0] = box.L[1] = box.L[2] = 50.0
    rc = lib.oxdna_pair_forces(term, None, None, None, 0, n, None, None, None, C.byref(box), None, None, None,
                               ntypes, C.cast(table, C.c_void_p), rcutsq, None, None)
    return rc, lib.oxdna_last_error(None).decode()


def _table(force, types=("A", "G", "C", "T")):
    table, rcut = force._pack(list(types))
    return table, (C.c_double * len(rcut))(*[r * r for r in rcut])


def test_nonuniform_tables_are_accepted(lib):
    nl = dna2.make_neighbor_list()
    ev = dna2.excluded_volume(nl)
    v = dict(ev.params[("A", "A")])
    v["eps"] = 3.0
    ev.params[("G", "T")] = v
    cx = dna2.coaxial_stacking(nl)
    w = dict(cx.params[("A", "A")])
    w["f2"] = (1.5 * w["f2"][0],) + tuple(w["f2"][1:])
    cx.params[("A", "C")] = w
    for f in (ev, cx):
        table, rcutsq = _table(f)
        rc, msg = _call(lib, f._term, 4, table, rcutsq)
        assert rc == 0, msg


def test_more_than_four_types_rejected(lib):
    nl = dna2.make_neighbor_list()
    ev = dna2.excluded_volume(nl, types=("A", "G", "C", "T", "X"))
    table, rcutsq = _table(ev, ("A", "G", "C", "T", "X"))
    rc, msg = _call(lib, ev._term, 5, table, rcutsq)
    assert rc != 0 and "ntypes must be 1..4" in msg


def test_bad_cutoffs_and_terms_rejected(lib):
    nl = dna2.make_neighbor_list()
    ev = dna2.excluded_volume(nl)
    table, rcutsq = _table(ev)
    rcutsq[5] = float("nan")
    rc, msg = _call(lib, ev._term, 4, table, rcutsq)
    assert rc != 0 and "cut-off" in msg
    table, rcutsq = _table(ev)
    rc, msg = _call(lib, _abi.TERM_BONDED, 4, table, rcutsq)
    assert rc != 0 and "term id" in msg
    assert np.isfinite(list(rcutsq)).all()
