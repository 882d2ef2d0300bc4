"""Stand-alone decks on the CPU: the reference's example decks read into
solver parameters, and the flowstate.dat line format against Fortran's own
formatted output.

The decks under tests/golden/decks are the input files of the reference's
Taylor_Green_Vortex and Riemann2D examples (data, copied unchanged).
"""
import os
import shutil
import subprocess

import pytest

from astr_gpu_amd import deck

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def test_taylor_green_deck():
    d = deck.read_input(deck.find_input(os.path.join(DECKS, "tgv")))
    assert d["flowtype"] == "tgv"
    assert (d["ia"], d["ja"], d["ka"]) == (128, 128, 128)
    assert (d["lihomo"], d["ljhomo"], d["lkhomo"]) == (True, True, True)
    assert d["nondimen"] and d["diffterm"] and d["lfilter"] and not d["lreadgrid"]
    assert d["alfa_filter"] == 0.49 and d["kcutoff"] == 48
    assert (d["ref_tem"], d["reynolds"], d["mach"]) == (273.15, 1600.0, 0.1)
    assert (d["conschm"], d["difschm"], d["rkscheme"]) == ("643c", "643c", "rk3")
    assert d["bctype"] == [1] * 6 and d["ninit"] == 0
    p = deck.params_from_deck(d)
    assert (p.im, p.jm, p.km) == (128, 128, 128)
    assert (p.npdci, p.npdcj, p.npdck) == (3, 3, 3)
    assert (p.convection, p.deriv_explicit, p.lfilter, p.diffterm) == (0, 0, 1, 1)
    assert (p.reynolds, p.mach, p.alfa_filter) == (1600.0, 0.1, 0.49)
    c = deck.read_controller(os.path.join(DECKS, "tgv", "datin", "controller"))
    assert (c["maxstep"], c["feqlist"], c["deltat"]) == (10, 1, 1e-3)


def test_riemann_deck():
    d = deck.read_input(deck.find_input(os.path.join(DECKS, "riem2d")))
    assert d["flowtype"] == "riem2d"
    assert (d["ia"], d["ja"], d["ka"]) == (400, 400, 0)
    assert d["bctype"] == [0, 0, 0, 0, 1, 1]
    p = deck.params_from_deck(d)
    # bounded i and j on one domain (npdc 4); k absent
    assert (p.npdci, p.npdcj, p.npdck) == (4, 4, 3)
    assert (p.convection, p.deriv_explicit) == (1, 1)
    assert (p.recon_schem, p.lchardecomp, p.lfilter, p.diffterm) == (3, 1, 0, 0)
    assert (p.bfacmpld, p.shkcrt, p.mach) == (0.7, 0.01, 10.0)
    c = deck.read_controller(os.path.join(DECKS, "riem2d", "datin", "controller"))
    assert (c["maxstep"], c["feqlist"], c["deltat"]) == (8000, 1, 1e-4)


def test_unbuilt_deck_settings_are_rejected(tmp_path):
    src = os.path.join(DECKS, "tgv", "datin", "input.tgv")
    for old, new, what in (("\n1\n1\n1 \n1\n1\n1\n", "\n11\n1\n1 \n1\n1\n1\n", "bctype"),
                           ("643c, 643c, rk3", "643c, 643c, rk4", "rk3"),
                           ("\ntgv\n", "\nchannel\n", "flow type")):
        text = open(src).read()
        assert old in text
        case = tmp_path / what.replace(" ", "_")
        (case / "datin").mkdir(parents=True)
        (case / "datin" / "input.tgv").write_text(text.replace(old, new))
        with pytest.raises(deck.DeckError, match=what):
            deck.params_from_deck(deck.read_input(deck.find_input(str(case))))


def test_deck_walls_and_their_temperatures(tmp_path):
    """readinput's bc lines (src/readwrite.F90:561-575): bctype 41 carries its
    face's wall temperature; the built conditions pass params_from_deck."""
    src = os.path.join(DECKS, "tgv", "datin", "input.tgv")
    text = open(src).read()
    old = "\n1\n1\n1 \n1\n1\n1\n"
    assert old in text
    case = tmp_path / "walls"
    (case / "datin").mkdir(parents=True)
    (case / "datin" / "input.tgv").write_text(
        text.replace(old, "\n41, 1.25\n42\n50 \n60\n1\n1\n"))
    d = deck.read_input(deck.find_input(str(case)))
    assert d["bctype"] == [41, 42, 50, 60, 1, 1]
    assert d["twall"][0] == 1.25 and d["twall"][1:] == [1.0] * 5
    deck.params_from_deck(d)


def test_fortran_e_known_strings():
    assert deck.fortran_e(0.125) == " 0.1250000000000E+00"
    assert deck.fortran_e(-1.5e-7) == "-0.1500000000000E-06"
    assert deck.fortran_e(0.0) == " 0.0000000000000E+00"
    assert deck.fortran_e(1.0) == " 0.1000000000000E+01"
    assert deck.fortran_e(12345.678) == " 0.1234567800000E+05"
    assert deck.fortran_e(1e120) == "*" * 20
    assert deck.flowstate_line(7, 0.028, [1.0]) == (
        "                    7  0.2800000000000E-01  0.1000000000000E+01")


VALUES = [0.125, -1.5e-7, 0.0, 1.0, 12345.678, 0.1 + 0.2, 2.0 / 3.0, -9.999999999999999e-1,
          3.141592653589793e-9, 6.02214076e23, 1.0000000000000497e-99]


@pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"),
                    reason="no Fortran compiler to print the reference format")
def test_flowstate_format_matches_fortran(tmp_path):
    """listinit's (1X,A20) header and listwrite's (1X,I20,n(1X,E20.13E2))
    lines, printed by a Fortran program, against the writer."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    lits = ", ".join(f"{v!r}".replace("e", "d") if "e" in repr(v) else f"{v!r}d0" for v in VALUES)
    n = len(VALUES)
    prog = f"""program p
  implicit none
  real(8) :: v({n}) = (/ {lits} /)
  write(*,"(3(1X,A20))") 'nstep', 'time', 'kenergy'
  write(*,"(1X,I20,{n}(1X,E20.13E2))") 123, v
end program p
"""
    (tmp_path / "p.f90").write_text(prog)
    exe = str(tmp_path / "p")
    subprocess.run([fc, "-o", exe, str(tmp_path / "p.f90")], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    assert out[0] == deck.flowstate_header(["nstep", "time", "kenergy"])
    assert out[1] == deck.flowstate_line(123, VALUES[0], VALUES[1:])
