"""Writes tests/golden/gaas_tau_fit.npz from the reference's GaAs example run
(Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/dBz/
only-eph/n2e16): the S_z series the run wrote (sz_elec_tot.out, time in a.u.
and S_z) and the outputs of that directory's fit.py (sfit.out: t in ps, the
data and the fitted curve; tau_t.dat: the time-resolved lifetimes).  These
pin the restated fit.

Run in a container that holds the reference tree:
  python tests/golden/make_gaas_tau.py /root/reference
"""
import os
import sys

import numpy as np

RUN = "Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/dBz/only-eph/n2e16"


def main(ref_root):
    run = os.path.join(ref_root, RUN)
    sz = np.loadtxt(os.path.join(run, "sz_elec_tot.out"))
    sfit = np.loadtxt(os.path.join(run, "sfit.out"))
    taut = np.loadtxt(os.path.join(run, "tau_t.dat"))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "gaas_tau_fit.npz"),
                        t_au=sz[:, 0], s=sz[:, 1], sfit_t=sfit[:, 0], sfit_s=sfit[:, 1], sfit_fit=sfit[:, 2],
                        taut_t=taut[:, 0], taut=taut[:, 1])
    print(f"{len(sz)} samples; sfit {sfit.shape}; tau_t {taut.shape}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
