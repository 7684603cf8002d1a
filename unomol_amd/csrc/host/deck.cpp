#include "deck.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>

namespace uno {

#include "basis_data.inc"

namespace {
constexpr double kBohrPerAng = 1.0 / 0.529177210903;
constexpr double kDeg = M_PI / 180.0;
}  // namespace

bool parse_deck(const std::string& text, Deck& d, std::string& err) {
    std::istringstream in(text);
    d = Deck();
    if (!(in >> d.nshell >> d.norb >> d.ncen >> d.maxl >> d.nelec >> d.maxits >> d.eps)) {
        err = "deck header is incomplete";
        return false;
    }
    if (!(in >> d.int_flag[0] >> d.int_flag[1] >> d.scf_flag[0] >> d.scf_flag[1] >> d.scf_flag[2] >>
          d.prt_flag[0] >> d.prt_flag[1] >> d.prt_flag[2])) {
        err = "deck flags are incomplete";
        return false;
    }
    if (d.nshell <= 0 || d.ncen <= 0 || d.norb <= 0) {
        err = "deck has no shells, centres or orbitals";
        return false;
    }
    if (d.maxl > 4) {
        // same limit and message as Basis.hpp:222
        err = "Angular Momentum is too large for present program";
        return false;
    }
    d.charge.resize(d.ncen);
    d.xyz.resize(3 * d.ncen);
    for (int c = 0; c < d.ncen; ++c)
        if (!(in >> d.charge[c] >> d.xyz[3 * c] >> d.xyz[3 * c + 1] >> d.xyz[3 * c + 2])) {
            err = "deck centre list is incomplete";
            return false;
        }
    int nfun = 0;
    for (int s = 0; s < d.nshell; ++s) {
        int npr, l, cen;
        if (!(in >> npr >> l >> cen) || npr <= 0 || l < 0 || l > 4 || cen < 0 || cen >= d.ncen) {
            err = "deck shell " + std::to_string(s) + " is malformed";
            return false;
        }
        d.sh_npr.push_back(npr);
        d.sh_l.push_back(l);
        d.sh_cen.push_back(cen);
        d.sh_poff.push_back((int)d.alpha.size());
        for (int p = 0; p < npr; ++p) {
            double a, c;
            if (!(in >> a >> c) || !(a > 0.0)) {
                err = "deck primitive list of shell " + std::to_string(s) + " is malformed";
                return false;
            }
            d.alpha.push_back(a);
            d.coef.push_back(c);
        }
        nfun += (l + 1) * (l + 2) / 2;
    }
    if (nfun != d.norb) {
        err = "deck norb " + std::to_string(d.norb) + " disagrees with its shells (" + std::to_string(nfun) + ")";
        return false;
    }
    return true;
}

bool read_deck(const std::string& path, Deck& d, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "could not open file " + path;   // Basis.hpp:187-189 wording
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_deck(ss.str(), d, err);
}

bool write_deck(const std::string& path, const Deck& d, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        err = "could not create " + path;
        return false;
    }
    std::fprintf(f, "%7d %7d %7d %3d\n%7d %7d\n  %.16e\n %d %d\n %d %d %d\n %d %d %d\n", d.nshell, d.norb, d.ncen,
                 d.maxl, d.nelec, d.maxits, d.eps, d.int_flag[0], d.int_flag[1], d.scf_flag[0], d.scf_flag[1],
                 d.scf_flag[2], d.prt_flag[0], d.prt_flag[1], d.prt_flag[2]);
    for (int c = 0; c < d.ncen; ++c)
        std::fprintf(f, " %18.10f %22.16f %22.16f %22.16f\n", d.charge[c], d.xyz[3 * c], d.xyz[3 * c + 1],
                     d.xyz[3 * c + 2]);
    for (int s = 0; s < d.nshell; ++s) {
        std::fprintf(f, " %3d %2d %6d\n", d.sh_npr[s], d.sh_l[s], d.sh_cen[s]);
        for (int p = 0; p < d.sh_npr[s]; ++p)
            std::fprintf(f, " %20.10f %20.10f\n", d.alpha[d.sh_poff[s] + p], d.coef[d.sh_poff[s] + p]);
    }
    std::fclose(f);
    return true;
}

bool build_deck(const std::vector<Atom>& atoms, const std::string& basis, int charge, Deck& d,
                std::string& err) {
    d = Deck();
    d.ncen = (int)atoms.size();
    int zsum = 0;
    for (const Atom& a : atoms) {
        d.charge.push_back((double)a.z);
        for (int x = 0; x < 3; ++x) d.xyz.push_back(a.r[x]);
        zsum += a.z;
        std::string want = basis;
        if (a.z == 1 && basis == "6-31G*") want = "6-31G";   // no polarisation on H
        const BasisEntry* e = nullptr;
        for (const BasisEntry& be : kBasisLib)
            if (be.z == a.z && want == be.name) e = &be;
        if (!e) {
            err = "no basis " + basis + " for Z=" + std::to_string(a.z);
            return false;
        }
        int cen = (int)d.charge.size() - 1;
        for (int s = 0; s < e->nshell; ++s) {
            const BasisShellDef& sd = e->sh[s];
            d.sh_l.push_back(sd.l);
            d.sh_npr.push_back(sd.npr);
            d.sh_cen.push_back(cen);
            d.sh_poff.push_back((int)d.alpha.size());
            for (int p = 0; p < sd.npr; ++p) {
                d.alpha.push_back(kBasisPrims[sd.first + p][0]);
                d.coef.push_back(kBasisPrims[sd.first + p][1]);
            }
            d.norb += (sd.l + 1) * (sd.l + 2) / 2;
            if (sd.l > d.maxl) d.maxl = sd.l;
        }
    }
    d.nshell = (int)d.sh_l.size();
    d.nelec = zsum - charge;
    if (d.nelec <= 0) {
        err = "no electrons";
        return false;
    }
    return true;
}

void water_monomer(std::vector<Atom>& atoms) {
    const double r = 1.809, th = 104.52 * kDeg;
    atoms.push_back({8, {0.0, 0.0, 0.0}});
    atoms.push_back({1, {0.0, r * std::sin(th / 2), r * std::cos(th / 2)}});
    atoms.push_back({1, {0.0, -r * std::sin(th / 2), r * std::cos(th / 2)}});
}

namespace {
// rotation matrix of a unit quaternion (w, x, y, z)
void quat_rot(double w, double x, double y, double z, double R[3][3]) {
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z);     R[0][2] = 2 * (x * z + w * y);
    R[1][0] = 2 * (x * y + w * z);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
    R[2][0] = 2 * (x * z - w * y);     R[2][1] = 2 * (y * z + w * x);     R[2][2] = 1 - 2 * (x * x + y * y);
}
}  // namespace

void water_cluster(int n, uint64_t seed, std::vector<Atom>& atoms) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    // cube edge for 1 g/cm^3: L^3 = n * 18.015 g/mol / N_A
    const double L = std::cbrt(n * 18.015 / 6.02214076e23) * 1e8;   // Angstrom
    const double dmin2 = 2.5 * 2.5;
    std::vector<double> ox;
    while ((int)ox.size() < 3 * n) {
        double p[3] = {U(rng) * L, U(rng) * L, U(rng) * L};
        bool ok = true;
        for (size_t k = 0; k < ox.size(); k += 3) {
            double dx = p[0] - ox[k], dy = p[1] - ox[k + 1], dz = p[2] - ox[k + 2];
            if (dx * dx + dy * dy + dz * dz < dmin2) { ok = false; break; }
        }
        if (ok) ox.insert(ox.end(), p, p + 3);
    }
    std::vector<Atom> mono;
    water_monomer(mono);
    for (int w = 0; w < n; ++w) {
        // Haar-uniform rotation from three uniforms (Shoemake)
        double u1 = U(rng), u2 = U(rng), u3 = U(rng);
        double qx = std::sqrt(1 - u1) * std::sin(2 * M_PI * u2), qy = std::sqrt(1 - u1) * std::cos(2 * M_PI * u2);
        double qz = std::sqrt(u1) * std::sin(2 * M_PI * u3), qw = std::sqrt(u1) * std::cos(2 * M_PI * u3);
        double R[3][3];
        quat_rot(qw, qx, qy, qz, R);
        for (const Atom& a : mono) {
            Atom b;
            b.z = a.z;
            for (int i = 0; i < 3; ++i)
                b.r[i] = ox[3 * w + i] * kBohrPerAng + R[i][0] * a.r[0] + R[i][1] * a.r[1] + R[i][2] * a.r[2];
            atoms.push_back(b);
        }
    }
}

namespace {
struct MolDef {
    const char* name;
};
const MolDef kMols[] = {{"H2"}, {"HF"}, {"H2O"}, {"NH3"}, {"CH4"}, {"N2"}, {"CO"}, {"CO2"}, {"C2H2"}};
}  // namespace

const char* small_molecule_name(int kind) { return kMols[kind % 9].name; }
int small_molecule_count() { return 9; }

void small_molecule(int kind, uint64_t seed, std::vector<Atom>& atoms) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    auto bond = [&](double ang) { return ang * kBohrPerAng * (1.0 + 0.05 * U(rng)); };
    auto angle = [&](double deg) { return (deg + 5.0 * U(rng)) * kDeg; };
    auto put = [&](int z, double x, double y, double zz) { atoms.push_back({z, {x, y, zz}}); };
    switch (kind % 9) {
        case 0: put(1, 0, 0, 0); put(1, 0, 0, bond(0.741)); break;
        case 1: put(9, 0, 0, 0); put(1, 0, 0, bond(0.917)); break;
        case 2: {
            double r1 = bond(0.9573), r2 = bond(0.9573), th = angle(104.52);
            put(8, 0, 0, 0); put(1, 0, 0, r1); put(1, 0, r2 * std::sin(th), r2 * std::cos(th));
            break;
        }
        case 3: {
            // three N-H bonds at a common polar angle with a distorted azimuth
            double pol = angle(68.0);
            put(7, 0, 0, 0);
            for (int k = 0; k < 3; ++k) {
                double r = bond(1.012), az = (120.0 * k + 5.0 * U(rng)) * kDeg;
                put(1, r * std::sin(pol) * std::cos(az), r * std::sin(pol) * std::sin(az), -r * std::cos(pol));
            }
            break;
        }
        case 4: {
            const double t[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
            put(6, 0, 0, 0);
            for (int k = 0; k < 4; ++k) {
                double r = bond(1.087) / std::sqrt(3.0);
                // small angular distortion: perturb each direction by up to ~5 deg
                double dx = 0.087 * U(rng), dy = 0.087 * U(rng);
                put(1, r * (t[k][0] + dx), r * (t[k][1] + dy), r * t[k][2]);
            }
            break;
        }
        case 5: put(7, 0, 0, 0); put(7, 0, 0, bond(1.098)); break;
        case 6: put(6, 0, 0, 0); put(8, 0, 0, bond(1.128)); break;
        case 7: {
            double r1 = bond(1.16), r2 = bond(1.16), th = angle(180.0 - 5.0);
            put(6, 0, 0, 0); put(8, 0, 0, r1); put(8, 0, r2 * std::sin(th), r2 * std::cos(th));
            break;
        }
        default: {
            double cc = bond(1.203), ch1 = bond(1.063), ch2 = bond(1.063), th = angle(180.0 - 5.0);
            put(6, 0, 0, 0); put(6, 0, 0, cc); put(1, 0, 0, -ch1);
            put(1, 0, ch2 * std::sin(th), cc - ch2 * std::cos(th));
            break;
        }
    }
}

}  // namespace uno
