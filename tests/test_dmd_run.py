"""The param.in front end of dmd_run (host code only: --dry-run stops before
any GPU call): the settings it resolves from the reference GaAs example's own
param.in and ldbd_data (tests/golden/gaas_ldbd), checked against what that
run printed, and its refusals of features the operator does not carry."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "gaas_ldbd")
RUN = os.path.join(ROOT, "denmat_dynm_development_amd", "dmd_run")


def _dry(folder):
    if not os.path.exists(RUN):
        pytest.fail(f"{RUN} is missing: build it with __graft_entry__.build()")
    r = subprocess.run([RUN, "--dry-run", str(folder)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def _num(text, key):
    m = re.search(r"\b" + re.escape(key) + r" = (\S+)", text)
    assert m, (key, text)
    return float(m.group(1))


def _gaas_folder(tmp_path, param_text=None):
    d = tmp_path / "run"
    (d / "ldbd_data").mkdir(parents=True)
    for f in ("ldbd_size.dat", "ldbd_R.dat", "ldbd_ek.bin", "ldbd_kvec.bin"):
        shutil.copy(os.path.join(GOLD, f), d / "ldbd_data" / f)
    if param_text is None:
        shutil.copy(os.path.join(GOLD, "param.in"), d / "param.in")
    else:
        (d / "param.in").write_text(param_text)
    return d


def test_dry_run_resolves_the_reference_gaas_settings(tmp_path):
    """The example's param.in (carrier_density = 2e16, Bzpert = 0.1, tstep =
    100, tend = 1e5, alg_ddmdteq = 1) with its ldbd_data: mu from the carrier
    density, the cell volume, the times and field in atomic units agree with
    what the reference run printed to `out` (lines 50-168, 236: T, volume,
    carrier_density, Bzpert, tstep, tend, mu)."""
    rc, out, err = _dry(_gaas_folder(tmp_path))
    assert rc == 0, err
    assert abs(_num(out, "mu") / 4.9841627e-02 - 1) < 2e-8  # "mu =  4.9841627e-02" (8 digits)
    assert abs(_num(out, "T") / 9.5004341e-04 - 1) < 1e-7
    assert abs(_num(out, "cell_size") / 304.795 - 1) < 2e-6  # "volume = 304.795"
    assert abs(_num(out, "tstep") / 4134.14 - 1) < 2e-6 and abs(_num(out, "tend") / 4.13414e+06 - 1) < 2e-6
    assert abs(_num(out, "tstep") / (100 * 41.34137417094296) - 1) < 1e-10  # fs of comm/Units.h
    m = re.search(r"Bpert = (\S+) (\S+) (\S+)", out)
    assert float(m.group(3)) == pytest.approx(4.25438e-07, rel=2e-6) and float(m.group(1)) == 0
    assert "alg_ddmdteq = 1" in out and "alg_picture = interaction" in out
    assert "alg_ode_method = rkf45" in out and "ode_epsabs = 1e-08" in out  # the reference's defaults
    assert "e-ph bands [4, 6)" in out and "nb_dm = 6 (nv_dm = 4)" in out
    assert 'key "pumpE" is not read' in out  # said, not silently dropped
    assert _num(out, "freq_compute_tau") == 10 and _num(out, "compute_tau_only") == 0  # freq_measure_ene = 10


def test_dry_run_parses_continuations_units_and_substeps(tmp_path):
    text = "\n".join([
        "# a comment = not a key",
        "alg_ddmdteq = 0",
        "tstep = 3 \\",
        "",
        "ode_hmax = 1",
        "tend = 30",
        "freq_measure = 2",
        "freq_measure_ene = 4",
        "compute_tau_only = 1",
        "alg_sparseP = 1",
        "alg_thr_sparseP = 1e-5",
        "alg_semiclassical = 1",
        "Bxpert = 0.5",
        "mu = 0.05",
        "scale_scatt = 0.7",
        "scale_eph = 0.9",
        "alg_picture = schrodinger",
        "alg_ode_method = rk4",
    ]) + "\n"
    rc, out, err = _dry(_gaas_folder(tmp_path, text))
    assert rc == 0, err
    fs = 41.34137417094296
    assert abs(_num(out, "tstep") / (3 * fs) - 1) < 1e-10 and "x 3 per tstep" in out  # RK4 step 1 fs x 3
    assert _num(out, "freq_measure") == 2
    assert _num(out, "freq_compute_tau") == 4 and _num(out, "compute_tau_only") == 1  # parameters.cpp:326
    assert _num(out, "alg_sparseP") == 1 and _num(out, "alg_Pin_is_sparse") == 0 and _num(out, "alg_thr_sparseP") == 1e-5
    assert _num(out, "alg_semiclassical") == 1
    assert abs(_num(out, "mu") / (0.05 / 27.21138505) - 1) < 1e-10
    assert _num(out, "scale_eph") == 0.9 and _num(out, "scale_ei") == 0.7
    assert abs(float(re.search(r"Bpert = (\S+)", out).group(1)) / (0.5 / 2.3505175675871e5) - 1) < 1e-10
    assert "alg_picture = schrodinger" in out


def test_dry_run_accepts_the_reference_eph_ei_ee_run(tmp_path):
    """The reference's linearized e-ph + e-i + e-e example (its own param.in,
    tests/golden/gaas_eph_ei_ee_param.in: alg_linearize = 1, eeMode =
    Pee_update, impMode = model_ionized, scrMode = medium with scrFormula left
    at its default RPA, the param.in's own "srcFormula" misspelling being an
    unread key there too) with the GaAs ldbd_data: every key is carried."""
    d = _gaas_folder(tmp_path)
    shutil.copy(os.path.join(os.path.dirname(GOLD), "gaas_eph_ei_ee_param.in"), d / "param.in")
    rc, out, err = _dry(d)
    assert rc == 0, err
    assert "alg_linearize = 1" in out and "eeMode = Pee_update" in out and "scrFormula = RPA" in out
    assert "impurities = 1" in out and "model_impurities = 1" in out
    assert 'key "srcFormula" is not read' in out


def test_dry_run_accepts_the_rebuild_inside_every_rhs(tmp_path):
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\neeMode = Pee_update\nscrMode = medium\n"
                                               "freq_update_ee_model = -1\n"))
    assert rc == 0, err
    assert "eeMode = Pee_update" in out


def test_dry_run_accepts_linearize_dPee_with_Pee_update(tmp_path):
    d = _gaas_folder(tmp_path, "tstep = 1\ntend = 10\nalg_linearize_dPee = 1\nalg_ddmdteq = 1\neeMode = Pee_update\n"
                               "freq_update_ee_model = 2\nscrMode = medium\n")
    rc, out, err = _dry(d)
    assert rc == 0, err
    assert "alg_linearize_dPee = 1" in out and "eeMode = Pee_update" in out


@pytest.mark.parametrize("line,msg", [
    ("eeMode = Pee_update\nscrMode = medium", "want to update Pee but set freq_update_ee_model as 0"),
    ("eeMode = Pee_fixed_at_eq", "scrMode should not be none if considering electron-electron"),
    ("alg_only_ee = 1\nalg_eph_enable = 0\nimpurity_density = 1e17\nimpMode = ab_neutral", "alg_only_ee is only possible"),
    ("eeMode = other", "eeMode must be none or Pee_fixed_at_eq"),
    ("alg_linearize_dPee = 1", "linearize_dPee must be used ddmdteq"),
    ("alg_linearize_dPee = 1\nalg_ddmdteq = 1", "only active when e-e scattering is active"),
    ("alg_linearize_dPee = 1\nalg_linearize = 1", "cannot be both true"),
    ("alg_linearize_dPee = 1\nalg_ddmdteq = 1\nalg_sparseP = 1", "linearization does not support sparse"),
    ("alg_linearize_dPee = 1\nalg_ddmdteq = 1\neeMode = Pee_fixed_at_eq\nscrMode = medium", "use eeMode = Pee_update"),
    ("impurity_density = 1e17\nimpMode = model_ionized", "scrMode should not be none"),
    ("impurity_density = 1e17\nscrMode = medium\nscrFormula = lindhard", "built for the static RPA"),
    ("impurity_density = 1e17\nscrMode = medium\nscrFormula = heg\ndynamic_screening = model", "built for the static RPA"),
    ("impurity_density = 1e17\nscrMode = medium\nscrFormula = debye\npartial_ionized = 1", "partially ionized"),
    ("impurity_density = 1e17\nscrMode = medium\nscrFormula = debye\nZ_impurity = 0", "Z is defined as a positive"),
    ("impurity_density = 1e17\nimpMode = other", "impMode must be model_ionized or ab_neutral"),
    ("scrMode = full", "scrMode must be none or medium"),
    ("scrMode = medium\nscrFormula = debye\ndynamic_screening = ppa", "static limit"),
    ("impurity_density = -1e17\nimpMode = ab_neutral\ncarrier_density = 1e17", "same sign"),
    ("alg_eph_need_hole = 1\nalg_eph_need_elec = 0", "file set"),
    ("pumpA0 = 0.002", "pump"),
    ("alg_ode_method = rk2", "rkf45, euler or rk4"),
    ("alg_Pin_is_sparse = 1\nalg_linearize = 1", "linearization for sparse P matrices is not implemented"),
    ("alg_thr_sparseP = -1", "alg_thr_sparseP"),
    ("alg_only_intravalley = 1\nalg_only_intervalley = 1", "cannot be true at the same time"),
    ("valley1 = 0.3 0.3 0", "at least you need provide positions of two valleys"),
    ("alg_eph_enable = 0", "alg_scatt_enable must be false"),
    ("alg_scatt_enable = 0", "alg_scatt_enable must be true"),
    ("alg_eph_enable = 0\nalg_only_eimp = 1", "only possible if impurity_density"),
    ("alg_only_eimp = 1\nimpurity_density = 1e17\nimpMode = ab_neutral", "Set alg_eph_enable = 0"),
    ("alg_sparseP = 1\nimpurity_density = 1e17\nimpMode = ab_neutral", "does not support sparse"),
    ("gfac_k_resolved = 0", "g factor should not have band dependence"),
    ("alg_eph_need_elec = 0\nalg_eph_need_hole = 0", "either alg_eph_need_elec or alg_eph_need_hole"),
])
def test_dry_run_refuses_what_the_operator_does_not_carry(tmp_path, line, msg):
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\n" + line + "\n"))
    assert rc != 0 and msg in err, (rc, err)


def test_teph_class_compiles_for_the_reference_complex_type(tmp_path):
    """Every member of electronphonon_mi355x, instantiated for a complex type
    laid out as the reference's (struct of two doubles, comm/scalar.h:67-69),
    and the members dm_dynamics uses (DensityMatrix_Dynamics.h:59-63,70,
    266-319) called with the reference's argument types."""
    src = tmp_path / "teph_check.cpp"
    src.write_text(r'''
#include "electronphonon_mi355x.h"
struct complex { double x, y; };
template class dmdrun::electronphonon_mi355x<complex>;
void driver_calls(dmdrun::electronphonon_mi355x<complex> *eph, complex **dm, complex **dm1, complex **ddmdt,
                  double **f_eq, double t) {
    eph->evolve_driver(t, dm, dm1, ddmdt);
    eph->compute_ddmdt_eq(f_eq);
    eph->reset_scatt(true, true, dm, nullptr, t, f_eq);
    int b = eph->bStart + eph->bEnd;
    bool no_models = eph->ee_model == nullptr && (eph->eimp.empty() || eph->eimp[0]->eimp_model == nullptr);
    (void)b; (void)no_models;
}
''')
    inc = [os.path.join(ROOT, "include"), os.path.join(ROOT, "denmat_dynm_development_amd", "host")]
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc[0], "-I", inc[1], str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_dry_run_scattering_switches(tmp_path):
    """alg_scatt_enable = 0 with alg_eph_enable = 0 (coherent dynamics only) and
    alg_eph_enable = 0 with an ab initio impurity (e-i only) are accepted."""
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\nalg_scatt_enable = 0\nalg_eph_enable = 0\n"))
    assert rc == 0, err
    assert _num(out, "alg_scatt_enable") == 0 and _num(out, "alg_eph_enable") == 0
    rc, out, err = _dry(_gaas_folder(tmp_path / "b", "tstep = 1\ntend = 10\nalg_eph_enable = 0\nalg_only_eimp = 1\n"
                                     "impurity_density = 1e17\nimpMode = ab_neutral\n"))
    assert rc == 0, err
    assert _num(out, "alg_only_eimp") == 1 and _num(out, "impurities") == 1


def test_dry_run_valleys_and_forbidden_transitions(tmp_path):
    """valley1, valley2, ... (crystal coordinates, parameters.cpp:330-350) and
    file_forbid_vtrans ("i j" per line, parameters.cpp:25-44)."""
    d = _gaas_folder(tmp_path, "tstep = 1\ntend = 10\nvalley1 = 0.3333333 0.3333333 0\n"
                               "valley2 = -0.3333333 -0.3333333 0\nfile_forbid_vtrans = forbid.txt\n"
                               "alg_only_intervalley = 1\n")
    (d / "forbid.txt").write_text("0 1\n")
    rc, out, err = _dry(d)
    assert rc == 0, err
    assert _num(out, "valleys") == 2 and _num(out, "forbidden_transitions") == 1 and _num(out, "alg_only_intervalley") == 1
    (d / "forbid.txt").write_text("0 2\n")
    rc, out, err = _dry(d)
    assert rc != 0 and "valley index" in err


def test_dry_run_excess_density_and_phenomenological_relaxation(tmp_path):
    """carrier_density_means_excess_density (electron.cpp:537-552: electrons
    minus holes over every band) gives the reference run's mu when the holes
    are negligible (GaAs gap at 300 K) and a mid-gap mu for zero excess;
    alg_phenom_relax reads tau_phenom in ps and the band range."""
    text = open(os.path.join(GOLD, "param.in")).read()
    rc, out, err = _dry(_gaas_folder(tmp_path, text + "\ncarrier_density_means_excess_density = 1\n"))
    assert rc == 0, err
    assert abs(_num(out, "mu") / 4.9841627e-02 - 1) < 2e-8
    rc, out, err = _dry(_gaas_folder(tmp_path / "z", "tstep = 1\ntend = 10\ncarrier_density = 0\n"
                                                   "carrier_density_means_excess_density = 1\n"))
    assert rc == 0, err
    e = np.fromfile(os.path.join(GOLD, "ldbd_ek.bin")).reshape(-1, 10)
    mu = _num(out, "mu")
    assert e[:, 5].max() < mu < e[:, 6].min()  # between the valence and conduction bands (nv = 6)
    rc, out, err = _dry(_gaas_folder(tmp_path / "p", "tstep = 1\ntend = 10\nalg_phenom_relax = 1\ntau_phenom = 2\n"
                                                   "bStart_tau = 1\nbEnd_tau = 5\n"))
    assert rc == 0, err
    assert abs(_num(out, "tau_phenom") / (2000 * 41.34137417094296) - 1) < 1e-10 and _num(out, "bEnd_tau") == 5
    rc, out, err = _dry(_gaas_folder(tmp_path / "q", "tstep = 1\ntend = 10\nalg_phenom_relax = 1\nbEnd_tau = 9\n"))
    assert rc != 0 and "bEnd_tau > nb_dm" in err


def test_dry_run_coherent_hamiltonian_inputs(tmp_path):
    """needL, alg_read_Bso, scale_Ez and alg_set_scv_zero (parameters.cpp:115,
    186, 295, 323) are read, not refused, and default to off."""
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\n"))
    assert rc == 0, err
    assert "alg_read_Bso = 0 needL = 0 scale_Ez = 0 alg_set_scv_zero = 0" in out
    rc, out, err = _dry(_gaas_folder(tmp_path / "b", "tstep = 1\ntend = 10\nneedL = 1\nalg_read_Bso = 1\n"
                                                   "scale_Ez = 0.25\nalg_set_scv_zero = 1\n"))
    assert rc == 0, err
    assert "alg_read_Bso = 1 needL = 1 scale_Ez = 0.25 alg_set_scv_zero = 1" in out
    assert "not read by this runner" not in out
    assert "gfac_normal_dist = 0 gfac_mean = 2 gfac_sigma = 0 gfac_cap = 0" in out


def test_dry_run_k_resolved_random_g_factors(tmp_path):
    """gfac_normal_dist with its mean, sigma and cap (parameters.cpp:179-183)."""
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\ngfac_normal_dist = 1\ngfac_mean = 2.1\n"
                                                "gfac_sigma = 0.3\ngfac_cap = 0.5\ngfac_k_resolved = 1\n"))
    assert rc == 0, err
    assert "gfac_normal_dist = 1 gfac_mean = 2.1 gfac_sigma = 0.3 gfac_cap = 0.5" in out
    assert "not read by this runner" not in out


def test_dry_run_model_ionized_impurity(tmp_path):
    """impMode = model_ionized (the default) with scrMode = medium and a
    |q|-only screening formula is accepted, with the reference's defaults
    (epsilon_background 1, detailBalance 0, degauss_eimp = degauss of
    ldbd_size.dat); it combines with alg_sparseP, which ab_neutral does not."""
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\nimpurity_density = 1e17\nscrMode = medium\n"
                                               "scrFormula = Bechstedt\nepsilon_background = 12.9\nalg_sparseP = 1\n"))
    assert rc == 0, err
    assert "scrMode = medium scrFormula = Bechstedt epsilon_background = 12.9 model_impurities = 1" in out
    degauss = float(open(os.path.join(GOLD, "ldbd_size.dat")).read().splitlines()[7].split()[0])
    assert _num(out, "degauss_eimp") == pytest.approx(degauss, rel=1e-12)
    assert "not read by this runner" not in out


def test_dry_run_carrier_set_selection(tmp_path):
    """One electronphonon per run with isHole = !need_elec (main.cpp:65,
    parameters.cpp:92-99): alg_eph_sepr_eh = 0 (both flags implied) and
    need_elec = need_hole = 1 take the conduction set of this ldbd_data."""
    for i, extra in enumerate(["alg_eph_sepr_eh = 0", "alg_eph_need_elec = 1\nalg_eph_need_hole = 1"]):
        rc, out, err = _dry(_gaas_folder(tmp_path / str(i), "tstep = 1\ntend = 10\n" + extra + "\n"))
        assert rc == 0, err


def test_dry_run_fixed_ee(tmp_path):
    """eeMode = Pee_fixed_at_eq with scrMode = medium is read with the
    reference's defaults (scale_ee = scale_scatt, degauss_ee = degauss)."""
    rc, out, err = _dry(_gaas_folder(tmp_path, "tstep = 1\ntend = 10\neeMode = Pee_fixed_at_eq\nscrMode = medium\n"
                                               "scrFormula = debye\nscale_scatt = 0.5\n"))
    assert rc == 0, err
    assert "eeMode = Pee_fixed_at_eq alg_only_ee = 0 scale_ee = 0.5" in out
    degauss = float(open(os.path.join(GOLD, "ldbd_size.dat")).read().splitlines()[7].split()[0])
    assert _num(out, "degauss_ee") == pytest.approx(degauss, rel=1e-12)
