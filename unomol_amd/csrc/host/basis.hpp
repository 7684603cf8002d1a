// Normalised shell model shared by the host code and the device-data builder.
// Normalisation follows Shell::normalize (Basis.hpp:56-75); Cartesian components are
// ordered lx descending then ly descending with per-component factors
// 1/sqrt(df(lx) df(ly) df(lz)) (AuxFunctions.hpp:35-64).
#pragma once
#include <string>
#include <vector>
#include "../../../include/unomol.h"

namespace uno {

struct Shell {
    int l, npr, cen;
    std::vector<double> a, c;   // exponents, normalised contraction coefficients
};

struct Basis {
    std::vector<Shell> sh;
    std::vector<double> xyz, q;   // centres (bohr), charges
    std::vector<int> off;         // first function of each shell
    int norb = 0, ncen = 0, maxl = 0;
};

constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// k-th Cartesian component of shell l: (lx, ly, lz)
void cart(int l, int k, int& lx, int& ly, int& lz);
// per-component normalisation factor, reference table semantics (exact for l <= 2)
double cart_norm(int lx, int ly, int lz);

bool make_basis(const uno_basis* b, Basis& out, std::string& err);

inline long pidx(int i, int j) { return i >= j ? (long)i * (i + 1) / 2 + j : (long)j * (j + 1) / 2 + i; }

// one-electron matrices (packed) and raw moment integrals (9 per packed element)
void one_electron(const Basis& B, double* S, double* T, double* H, double* M);

}  // namespace uno
