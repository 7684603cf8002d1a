// Deck (patin.dat) model, reader/writer and the seeded molecule / water-cluster
// generator.  The deck layout is the reference's input format (Basis.hpp:181-255):
//   nshell norb ncen maxl / nelec maxits / eps / int_flag[2] / scf_flag[3] /
//   prt_flag[3] / ncen lines "Z x y z" (bohr) / nshell shells "npr l cen" + npr
//   lines "alpha coef" (raw, un-normalised coefficients).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace uno {

struct Deck {
    int nshell = 0, norb = 0, ncen = 0, maxl = 0, nelec = 0, maxits = 300;
    double eps = 1e-10;
    int int_flag[2] = {0, 0};
    int scf_flag[3] = {2, 0, 0};   // convergence type (0 E, 1 P, 2 both), accel, guess
    int prt_flag[3] = {0, 0, 0};
    std::vector<double> charge, xyz;                 // [ncen], [3*ncen] bohr
    std::vector<int> sh_l, sh_npr, sh_cen, sh_poff;  // [nshell]
    std::vector<double> alpha, coef;                 // raw primitives
};

struct Atom {
    int z;
    double r[3];   // bohr
};

// read/write the deck text format; return false and fill err on failure
bool read_deck(const std::string& path, Deck& d, std::string& err);
bool parse_deck(const std::string& text, Deck& d, std::string& err);
bool write_deck(const std::string& path, const Deck& d, std::string& err);

// deck for a list of atoms in a named basis ("STO-3G", "6-31G", "6-31G*", "6-31G**");
// H uses the unpolarised entry in 6-31G*.  nelec = sum(Z) - charge.
bool build_deck(const std::vector<Atom>& atoms, const std::string& basis, int charge, Deck& d,
                std::string& err);

// rigid water monomer: r(OH) = 1.809 bohr, HOH = 104.52 deg, O at the origin,
// hydrogens at (0, +-r sin(theta/2), r cos(theta/2))
void water_monomer(std::vector<Atom>& atoms);

// n rigid waters: oxygens uniform in a cube sized for 1 g/cm^3, O-O >= 2.5 Angstrom
// (rejection), Haar-uniform orientations (unit quaternions), std::mt19937_64(seed).
void water_cluster(int n, uint64_t seed, std::vector<Atom>& atoms);

// seeded distortion (bond lengths +-5%, angles +-5 deg) of a small molecule from
// the pool {H2, HF, H2O, NH3, CH4, N2, CO, CO2, C2H2}; kind indexes the pool.
void small_molecule(int kind, uint64_t seed, std::vector<Atom>& atoms);
const char* small_molecule_name(int kind);
int small_molecule_count();

}  // namespace uno
