"""Copies the small ldbd_data files of the reference's GaAs example run
(Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/
dBz/only-eph/n2e16) into tests/golden/gaas_ldbd/, unchanged, as inputs of the
reader tests: ldbd_size.dat, ldbd_R.dat, ldbd_ek.bin, ldbd_kvec.bin,
ldbd_imsig.bin, restart/time_restart.dat and the run's param.in (about 115 KB), and
the param.in of the example's linearized e-ph + e-i + e-e run as
tests/golden/gaas_eph_ei_ee_param.in.  The run's P and
k-pair files were never shipped with the example; ldbd_smat.bin (1.6 MB) and
denmat_restart.bin stay out, their contents enter gaas_spin_restart.npz.

Run in a container that holds the reference tree:
  python tests/golden/make_gaas_ldbd.py /root/reference
"""
import os
import shutil
import sys

RUN = "Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/dBz/only-eph/n2e16"
FILES = ["ldbd_data/ldbd_size.dat", "ldbd_data/ldbd_R.dat", "ldbd_data/ldbd_ek.bin", "ldbd_data/ldbd_kvec.bin",
         "ldbd_data/ldbd_imsig.bin", "restart/time_restart.dat", "param.in"]


# the linearized e-ph + e-i + e-e run of the same example (its param.in only)
RUN_EE = "Examples/RealTime_Example_GaAs_20230105/ld-dmd/t300/mult8_gauss0.005/dBz/eph-ei-ee/n2e16_ni2e16"


def main(ref_root):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gaas_ldbd")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(ref_root, RUN, f), os.path.join(out, os.path.basename(f)))
        print(f, os.path.getsize(os.path.join(out, os.path.basename(f))))
    dst = os.path.join(os.path.dirname(out), "gaas_eph_ei_ee_param.in")
    shutil.copyfile(os.path.join(ref_root, RUN_EE, "param.in"), dst)
    print(RUN_EE + "/param.in", os.path.getsize(dst))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
