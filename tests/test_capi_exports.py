"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/lepic.h (no compute calls: there is no GPU here)."""
import os
import re

import pytest

from jd_lepic_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lepic.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(lepic_\w+)\s*\(", text, re.M)))


def test_header_and_bindings_agree():
    assert declared_symbols() == sorted(capi.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="library not built")
def test_library_exports_every_symbol():
    lib = capi.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_tables_match_oracle_ordering():
    import numpy as np
    import oracle as O
    from jd_lepic_amd import reactions as RX, tables
    for chem in (RX.argon(7.2429e19), RX.hydrogen(1e20)):
        t = tables.ordering(chem)
        o = O.ordering([dict(col_info=r.col_info(), Eth=r.Eth, E=r.E, S=r.S) for r in chem.reactions],
                       chem.masses(), chem.ntype_all)
        assert np.array_equal(t["sig_Er"], o["sig_Er"])
        assert np.array_equal(t["sig"], o["sig"])
        nt = chem.ntype_all
        assert np.array_equal(t["p_ncol"][1:], o["p_ncol"][1:nt + 1])
        for p in range(1, nt + 1):
            assert np.array_equal(t["sigv_mx"][p, :t["ncol"]], o["sigv_mx"][p][:t["ncol"]])
            assert np.array_equal(t["sig_list"][p, :t["p_ncol"][p]], o["sig_list"][p][1:t["p_ncol"][p] + 1])


def test_reaction_file_round_trip():
    import numpy as np
    from jd_lepic_amd import reactions as RX
    chem = RX.argon(7.2429e19)
    text = RX.write(chem, 7.2429e19)
    back = RX.parse(text, 7.2429e19)
    assert [s.name for s in back.species] == [s.name for s in chem.species]
    for a, b in zip(chem.reactions, back.reactions):
        assert a.col_info() == b.col_info()
        assert np.allclose(a.E, b.E, rtol=1e-15) and np.allclose(a.S, b.S, rtol=1e-15)
