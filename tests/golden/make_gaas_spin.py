"""Writes tests/golden/gaas_spin_restart.npz from the reference's GaAs example
run (Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/
dBz/only-eph/n2e16): the final density matrix (restart/denmat_restart.bin),
the S_z matrices and band energies of the density-matrix window
(ldbd_data/ldbd_smat.bin, ldbd_ek.bin), the run's T, mu, nk_full and lattice
vectors (ldbd_size.dat, ldbd_R.dat, the mu printed in `out`), and the last
line of sz_elec_tot.out, which the reference wrote at the restart time.

Only the data the electron S_z observable reads are kept: for every k the
rows i in [nv_dm, nb_dm) of S_z and the columns i of rho.  File formats:
ldbd_ek.bin nk x nb float64 (electron.cpp:100-175), ldbd_smat.bin
nk x 3 x nb_dm^2 complex128 (electron.cpp:284-314), denmat_restart.bin
nk x nb_dm^2 complex128 (DenMat.cpp:301-334).

Run in a container that holds the reference tree:
  python tests/golden/make_gaas_spin.py /root/reference
"""
import os
import sys

import numpy as np

RUN = "Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/dBz/only-eph/n2e16"


def main(ref_root):
    run = os.path.join(ref_root, RUN)
    size = open(os.path.join(run, "ldbd_data/ldbd_size.dat")).read().split("\n")
    nb, nv, b0_dm, b1_dm = (int(x) for x in size[1].split()[:4])
    nk_full, nk = float(size[2].split()[0]), int(size[2].split()[1])
    T = float(size[5].split()[0])  # line 1 is a comment, line 6 the temperature
    nb_dm, nv_dm = b1_dm - b0_dm, nv - b0_dm
    # the run's mu: the value the carrier-density solve settled on, printed in `out`
    mu = None
    for line in open(os.path.join(run, "out")):
        if line.startswith("isHole = 0 mu0 ="):
            mu = float(line.split("mu =")[-1].strip().rstrip(":"))
    # ldbd_R.dat: the dimension on line 1, then the lattice vectors in bohr (lattice.h:48-56)
    R = np.loadtxt(os.path.join(run, "ldbd_data/ldbd_R.dat"), skiprows=1)
    e = np.fromfile(os.path.join(run, "ldbd_data/ldbd_ek.bin"), np.float64).reshape(nk, nb)
    s = np.fromfile(os.path.join(run, "ldbd_data/ldbd_smat.bin"), np.complex128).reshape(nk, 3, nb_dm, nb_dm)
    dm = np.fromfile(os.path.join(run, "restart/denmat_restart.bin"), np.complex128).reshape(nk, nb_dm, nb_dm)
    t_restart = float(open(os.path.join(run, "restart/time_restart.dat")).read())
    last = [float(x) for x in open(os.path.join(run, "sz_elec_tot.out")).read().strip().split("\n")[-1].split()]
    assert abs(last[0] - t_restart) < 1e-6 * t_restart
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "gaas_spin_restart.npz"),
        e_dm=e[:, b0_dm:b1_dm], sz_rows=s[:, 2, nv_dm:, :], dm_cols=dm[:, :, nv_dm:],
        nv_dm=nv_dm, nb_dm=nb_dm, T=T, mu=mu, nk_full=nk_full, R=R,
        t=last[0], sz=last[1], sz_amp=last[2])
    print(f"nk={nk} nb_dm={nb_dm} nv_dm={nv_dm} T={T} mu={mu} t={last[0]} sz={last[1]} sz_amp={last[2]}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
