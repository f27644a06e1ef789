"""Extracts the screened Coulomb table that the reference's GaAs e-ph + e-i
example run printed (Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/
mult8_gauss0.005/dBz/eph-ei/n2e16_ni2e16/ldbd_vq.out) into
tests/golden/gaas_vq.npz: the three columns |q|^2, |q_scr|^2 and v(q) of its
6807 rows, and the numbers of its two header lines (n_free in cm^-3, the Debye
q^2, the smallest nonzero |q|^2, k_F^2 and the Thomas-Fermi q^2).  The run's
ldbd_data has the k mesh and lattice of tests/golden/gaas_ldbd (same
ldbd_kvec.bin and ldbd_R.dat); its param.in sets epsilon_background = 12.9 at
300 K.

Run in a container that holds the reference tree:
  python tests/golden/make_gaas_vq.py /root/reference
"""
import os
import re
import sys

import numpy as np

RUN = "Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/dBz/eph-ei/n2e16_ni2e16"


def main(ref_root):
    path = os.path.join(ref_root, RUN, "ldbd_vq.out")
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[0] + " " + lines[1]
    num = r"\s*([-+]?[0-9.]+(?:[eE][-+]?[0-9]+)?)"
    key = {k: float(re.search(re.escape(k) + num, head).group(1))
           for k in ("nfree =", "qscr2_debye =", "|qmin|^2 =", "kF2 =", "qTF2 =")}
    rows = np.array([[float(x) for x in l.split()] for l in lines[3:] if l.strip()])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gaas_vq.npz")
    np.savez_compressed(out, q2=rows[:, 0], qscr2=rows[:, 1], vq=rows[:, 2], nfree_cm3=key["nfree ="],
                        qscr2_debye=key["qscr2_debye ="], qmin2=key["|qmin|^2 ="], kF2=key["kF2 ="],
                        qTF2=key["qTF2 ="])
    print(out, rows.shape, key, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
