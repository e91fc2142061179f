// Fermionic lattice presets (reference: include/pomerol/LatticePresets.hpp,
// src/pomerol/LatticePresets.cpp:46-126,140-147,402-409).  Index tuple: (site label, orbital,
// spin) with spin numbering undef=-1, down=0, up=1 (LatticePresets.hpp:31-35).
#pragma once

#include "Operators.hpp"

#include <string>

namespace Pomerol {
namespace LatticePresets {

enum spin : int { undef = -1, down = 0, up = 1 };

inline std::ostream& operator<<(std::ostream& os, spin s) {
    return os << (s == up ? "up" : (s == down ? "dn" : ""));
}

using RealExpr = Operators::expression<RealType, std::string, unsigned short, spin>;
using ComplexExpr = Operators::expression<ComplexType, std::string, unsigned short, spin>;

inline RealExpr c_dag(std::string const& label, unsigned short orbital, spin s) {
    return Operators::c_dag<RealType>(label, orbital, s);
}
inline RealExpr c(std::string const& label, unsigned short orbital, spin s) {
    return Operators::c<RealType>(label, orbital, s);
}
inline RealExpr n(std::string const& label, unsigned short orbital, spin s) {
    return Operators::n<RealType>(label, orbital, s);
}

/// eps * n (LatticePresets.cpp:46-48)
inline RealExpr Level(std::string const& Label, RealType Eps, unsigned short Orbital, spin Spin) {
    return Eps * n(Label, Orbital, Spin);
}
/// eps * (n_up + n_down) for every orbital (LatticePresets.cpp:53-61)
inline RealExpr Level(std::string const& Label, RealType Eps, unsigned short NOrbitals = 1) {
    RealExpr res;
    for(unsigned short o = 0; o < NOrbitals; ++o) {
        res += Level(Label, Eps, o, up);
        res += Level(Label, Eps, o, down);
    }
    return res;
}

/// t c^+_1 c_2 + h.c. (LatticePresets.cpp:75-84)
inline RealExpr Hopping(std::string const& Label1, std::string const& Label2, RealType t, unsigned short Orbital1,
                        unsigned short Orbital2, spin Spin1, spin Spin2) {
    return (t * c_dag(Label1, Orbital1, Spin1) * c(Label2, Orbital2, Spin2)) + Operators::hc;
}
inline RealExpr Hopping(std::string const& Label1, std::string const& Label2, RealType t, unsigned short Orbital,
                        spin Spin) {
    return Hopping(Label1, Label2, t, Orbital, Orbital, Spin, Spin);
}
inline RealExpr Hopping(std::string const& Label1, std::string const& Label2, RealType t, unsigned short Orbital1,
                        unsigned short Orbital2) {
    return Hopping(Label1, Label2, t, Orbital1, Orbital2, up, up) + Hopping(Label1, Label2, t, Orbital1, Orbital2, down, down);
}
/// both spins of every orbital (LatticePresets.cpp:117-126)
inline RealExpr Hopping(std::string const& Label1, std::string const& Label2, RealType t, unsigned short NOrbitals = 1) {
    RealExpr res;
    for(unsigned short o = 0; o < NOrbitals; ++o) {
        res += Hopping(Label1, Label2, t, o, up);
        res += Hopping(Label1, Label2, t, o, down);
    }
    return res;
}

/// H (n_up - n_down) (LatticePresets.cpp:140-147)
inline RealExpr Magnetization(std::string const& Label, RealType H, unsigned short NOrbitals = 1) {
    RealExpr res;
    for(unsigned short o = 0; o < NOrbitals; ++o) {
        res += Level(Label, H, o, up);
        res += Level(Label, -H, o, down);
    }
    return res;
}

/// Delta c^+_{1 o1 s1} c^+_{2 o2 s2} + h.c. (LatticePresets.cpp:159-211, real amplitudes)
inline RealExpr Pairing(std::string const& Label1, std::string const& Label2, RealType Delta, unsigned short Orbital1,
                        unsigned short Orbital2, spin Spin1, spin Spin2) {
    return (Delta * c_dag(Label1, Orbital1, Spin1) * c_dag(Label2, Orbital2, Spin2)) + Operators::hc;
}
inline RealExpr Pairing(std::string const& Label1, std::string const& Label2, RealType Delta, unsigned short Orbital1,
                        unsigned short Orbital2) {
    return Pairing(Label1, Label2, Delta, Orbital1, Orbital2, up, down);
}
/// Delta c^+_{o up} c^+_{o dn} + h.c. on every orbital of one site
inline RealExpr Pairing(std::string const& Label, RealType Delta, unsigned short NOrbitals = 1) {
    RealExpr res;
    for(unsigned short o = 0; o < NOrbitals; ++o)
        res += Pairing(Label, Label, Delta, o, o);
    return res;
}

/// U n_{o1 s1} n_{o2 s2}
inline RealExpr NupNdown(std::string const& Label, RealType U, unsigned short Orbital1, unsigned short Orbital2,
                         spin Spin1, spin Spin2) {
    return U * n(Label, Orbital1, Spin1) * n(Label, Orbital2, Spin2);
}

/// eps (n_up + n_down) + U n_up n_down per orbital (LatticePresets.cpp:402-409)
inline RealExpr CoulombS(std::string const& Label, RealType U, RealType Eps, unsigned short NOrbitals = 1) {
    RealExpr res;
    for(unsigned short o = 0; o < NOrbitals; ++o)
        res += Level(Label, Eps, o, up) + Level(Label, Eps, o, down) + NupNdown(Label, U, o, o, up, down);
    return res;
}

} // namespace LatticePresets
} // namespace Pomerol
