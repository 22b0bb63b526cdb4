"""Host-side checks that need no GPU: the C-ABI libraries load and export every
declared entry point, and the input grammar rejects what the reference rejects
(ReadInput.h error paths, odd grids without a Gamma point)."""
import ctypes as C
import os
import re

import pytest

import helpers as Hh
from helpers import N


def _declared(header):
    text = open(os.path.join(Hh.ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(edus_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("edus.h", "libedus_hip.so"), ("edus_host.h", "libedus_host.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) >= 8
    dll = C.CDLL(os.path.join(N.LIBDIR, lib))
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def _write(tmp_path, text):
    p = tmp_path / "input.txt"
    p.write_text(text)
    return str(p)


BASE = open(os.path.join(Hh.AB5, "input.txt")).read()


@pytest.mark.parametrize("mutate,message", [
    (lambda s: s.replace("nkPT{", "nkPT"), "Can't find the character {"),
    (lambda s: s.replace("corehole 0.00396892 au", "corehole 0.00396892 furlong"), "Can't understand units furlong"),
    (lambda s: s.replace("corehole 0.00396892 au", "sideways 3"), "Unknow key sideways"),
    (lambda s: s.replace("it_resolution 5", "it_resolution 5\nSpectrum"), "option Spectrum not available"),
    (lambda s: s.replace("1 70 70", "1 71 71"), "must be even"),
    (lambda s: s.replace("nkPT{", "nkPT read {"), "Cannot read the k points of nkPT read"),
    (lambda s: s.replace("Haldane_CoreBN_Nedge", "GaSe_kp2"), "not supported by this library"),
])
def test_input_errors(tmp_path, mutate, message):
    with pytest.raises(N.EdusError, match=re.escape(message)):
        N.Setup(_write(tmp_path, mutate(BASE)))


def test_units_and_defaults(tmp_path):
    text = BASE.replace("dt  0.1 au", "dt  0.0024188843265050 fs").replace("it_resolution 5", "it_resolution 7")
    s = N.Setup(_write(tmp_path, text))
    assert abs(s.info["dt"] - 0.1) < 1e-12
    assert s.info["it_resolution"] == 7
    # t_fin <= 0 falls back to the pump duration (runge_kutta_init.cpp:356-370)
    text2 = BASE.replace("t_fin 1650.36549334624544 au", "t_fin -5 au")
    s2 = N.Setup(_write(tmp_path, text2))
    pump = s2.pulse(0)
    assert s2.info["t_fin"] == int(pump["ncycle"] * pump["period"] / 0.1) * 0.1


def test_missing_wannier_file(tmp_path):
    text = ("missing_seed\nunitcell angstrom {\n3.35 0 0\n0 2.165 1.25\n0 2.165 -1.25\n}\n"
            "tdse{\nWannier\ndt 0.1 au\nt_fin 10 au\n}\nbands{\nvalence 1\nconduction 1\n}\nnkPT{\n1 8 8\n}\n"
            "laserpump sin2 {\ncycles 3\nwavelength 400 nm\nintensity 1e11 wcm2\npolarization 0 0 1\n}\n")
    with pytest.raises(N.EdusError, match="missing_seed_tb.dat"):
        N.Setup(_write(tmp_path, text))


def test_print_population_options(tmp_path):
    """PrintPopulation t0 t1 dt (a.u., ReadInput.h:1022-1028) and its defaults
    (variables.cpp:84-86: 0, 100 and 20 times the fs->a.u. factor as written)."""
    line = [x for x in BASE.split("\n") if x.startswith("PrintPopulation")][0]
    # the reference's own AB5 input: a window starting after the run (no snapshots)
    i = N.Setup(_write(tmp_path, BASE)).info
    assert (i["print_flags"], i["print_start"], i["print_end"], i["print_step"]) == (
        1, 45888.924403583114, 82.68274667312272, 0.8268274667312273)
    i = N.Setup(_write(tmp_path, BASE.replace(line, "PrintPopulation"))).info
    assert (i["print_flags"], i["print_start"], i["print_end"], i["print_step"]) == (
        1, 0.0, 100.0 * 0.02418884326505, 20.0 * 0.02418884326505)
    i = N.Setup(_write(tmp_path, BASE.replace(line, ""))).info
    assert i["print_flags"] == 0


def test_batch_rejects_multistep_problems(tmp_path):
    """Batches run RK4 problems; the check happens before any device call."""
    import helpers as Hh
    p = Hh.coulomb_input(str(tmp_path), nk=6, stepper="Adams_Bashforth 3\n")
    s = N.Setup(p)
    with pytest.raises(RuntimeError, match="RK4|Coulomb"):
        N.Batch([s.problem()])
