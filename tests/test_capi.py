"""The C-ABI library loads without a GPU and exports every entry point that
include/amcfd.h declares; deck parsing (host logic) and its error codes."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "amcfd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(amcfd_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(amcfd):
    lib = ctypes.CDLL(amcfd.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(amcfd.EXPORTS)


def test_version_string(amcfd):
    assert "gfx950" in amcfd.version()


def test_deck_params_derived_constants(amcfd, case_dir):
    import decks
    d = amcfd.Deck(case_dir("plate120"))
    p = d.params
    assert (p.ni, p.nj, p.nk) == (122, 50, 34)
    assert p.nk_active0 == p.nk
    assert p.discretization == 1 and p.navierstokes == 1 and p.steady == 0
    # DomainManager.cpp:252-254
    cpavg = (p.acpa * p.tsolid + p.acpb + p.acpl) * 0.5
    assert p.hlcal == p.hsmelt + cpavg * (p.tliquid - p.tsolid)
    assert p.hlatnt == p.hlfriz - p.hlcal
    assert abs(p.hlatnt - decks.LATENT) < 1e-6
    assert (p.alaspow, p.alaseta, p.alasrb, p.alasfact) == (200.0, 0.3, 50e-6, 2.0)
    assert (p.avolpow, p.heatthick) == (0.0, 1e-5)
    assert p.maxit == 100 and p.nonlintol == 1e-6 and p.delt == 5e-6


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_deck_errors(amcfd, tmp_path, case_dir):
    good = open(case_dir("tiny")).read()
    # a blank line (the reference throws std::out_of_range from line.at(0))
    bad = _write(tmp_path, "blank.k", good.replace("*PARAMETERS", "\n*PARAMETERS"))
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.Deck(bad)
    assert e.value.code == -3
    # no *MESH_FILE (the reference exits with an error, Mesh.cpp:580-586)
    nomesh = _write(tmp_path, "nomesh.k", good.split("*MESH_FILE")[0])
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.Deck(nomesh)
    assert e.value.code == -3
    # mesh file missing (Mesh.cpp:485-493)
    missing = _write(tmp_path, "missing.k", good.split("*MESH_FILE")[0] + "*MESH_FILE\n/nonexistent/mesh.k\n")
    with pytest.raises(amcfd.AmcfdError) as e:
        amcfd.Deck(missing)
    assert e.value.code == -2
    # short value lists, which the reference reads past (DomainManager.cpp:123-198)
    for kw in ("*SOLVE_SPECIES\n1.0 5e-9 0.5", "*TRACK_ENERGY_BASED_FREE_SURFACE\n-1e6 1e6 1.8 20 1e-6 30 1e-16",
               "*OUTPUT_SECTION\n0 0 0 1 1 1"):
        t = _write(tmp_path, "short.k", good.replace("*MESH_FILE", kw + "\n*MESH_FILE"))
        with pytest.raises(amcfd.AmcfdError) as e:
            amcfd.Deck(t)
        assert e.value.code == -3, kw
    # a moving zone needs a zone on either side in x and y (the reference
    # writes past its zone arrays otherwise); this grid has two y zones
    for kw in ("*ACTIVATE_MOVING_MESH\n2 2", "*ACTIVATE_MOVING_MESH\n2"):
        t = _write(tmp_path, "mm.k", good.replace("*MESH_FILE", kw + "\n*MESH_FILE"))
        with pytest.raises(amcfd.AmcfdError) as e:
            amcfd.Deck(t)
        assert e.value.code == -3


def test_species_and_solidification_keywords(amcfd, case_dir, tmp_path):
    p = amcfd.Deck(case_dir("tiny_species")).params
    assert (p.species, p.solidification) == (1, 1)
    assert (p.initconcentration, p.massdiff, p.kp, p.urfc) == (1.0, 5e-9, 0.5, 0.7)
    # the solidification flag line, like *FLUID_FLOW_OFF, leaves the open
    # block open: values after it still belong to the block (Mesh.cpp:53-62)
    good = open(case_dir("tiny")).read()
    t = good.replace("  0.8 0.8 0.8 0.8 0.8", "  0.7 0.6\n*CALCULATE_SOLIDIFICATION_PARAMETERS\n  0.5 0.4 0.3")
    q = amcfd.Deck(_write(tmp_path, "s.k", t)).params
    assert q.solidification == 1 and (q.urfu, q.urfv, q.urfw, q.urfp, q.urfh) == (0.7, 0.6, 0.5, 0.4, 0.3)
    assert amcfd.Deck(case_dir("tiny")).params.species == 0


def test_moving_mesh_keyword(amcfd, case_dir):
    p = amcfd.Deck(case_dir("tiny_move")).params
    assert (p.movingmesh, p.xmoveindex, p.ymoveindex) == (1, 1, 1)
    assert p.udtoolpath == 1 and p.n_toolpath == 5
    assert amcfd.Deck(case_dir("tiny")).params.movingmesh == 0


def test_deck_grammar_details(amcfd, tmp_path, case_dir):
    good = open(case_dir("tiny")).read()
    # values may span lines; comments and unknown keywords reset the block
    t = good.replace("  0.8 0.8 0.8 0.8 0.8", "  0.7 0.6\n$ comment\n  0.5 0.4 0.3")
    t = t.replace("*SCALAR_OUT", "*SOME_UNKNOWN_KEYWORD\n1 2 3\n*SCALAR_OUT")
    d = amcfd.Deck(_write(tmp_path, "g.k", t))
    p = d.params
    assert (p.urfu, p.urfv, p.urfw, p.urfp, p.urfh) == (0.7, 0.6, 0.5, 0.4, 0.3)
    # *FLUID_FLOW_OFF and upwind default
    t2 = good.replace("*SELECT_DISCRETIZATION\npower_law\n", "*FLUID_FLOW_OFF\n")
    p2 = amcfd.Deck(_write(tmp_path, "g2.k", t2)).params
    assert p2.navierstokes == 0 and p2.discretization == 0


def test_longest_lines_accepted_one_more_refused(amcfd, case_dir):
    """x-lines of 1024 interior cells are the line solvers' limit; a deck with
    1025 is refused with AMCFD_E_CAPACITY before the device is touched."""
    import ctypes
    deck = amcfd.Deck(case_dir("tiny_long"))
    assert deck.params.ni - 2 == 1024
    p = amcfd.Params.from_buffer_copy(deck.params)
    p.ni = 1027
    h = ctypes.c_void_p()
    rc = amcfd.load_library().amcfd_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert rc == -7 and not h.value   # AMCFD_E_CAPACITY
    assert b"1024" in amcfd.load_library().amcfd_last_error()
