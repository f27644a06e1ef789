// units.h -- the measurement units param.in values are given in, in Hartree
// atomic units, derived from the same base values as the reference's
// src_v4.5.7/comm/Units.h:31-72 so that "tstep = 100" means the same a.u.
// time in both programs.
#pragma once
#include <cmath>

namespace dmdrun {
namespace units {
constexpr double eV = 1 / 27.21138505;
constexpr double Joule = 1 / 4.35974434e-18;
constexpr double Kelvin = 1.3806488e-23 * Joule;
constexpr double Angstrom = 1 / 0.5291772;
constexpr double meter = 1e10 * Angstrom;
constexpr double kg = 1. / 9.10938291e-31;
constexpr double Newton = Joule / meter;
inline double sec() { return std::sqrt((kg * meter) / Newton); }
inline double fs() { return sec() * 1.0e-15; }
inline double ps() { return 1000 * fs(); }
constexpr double bohr2cm = 5.291772109038e-9;
constexpr double Tesla2au = 1. / 2.3505175675871e5;
}  // namespace units
}  // namespace dmdrun
