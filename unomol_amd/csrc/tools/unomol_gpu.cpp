// Drop-in replacement of the reference program (Unomol.cc:8-26): reads patin.dat in
// the working directory, runs RHF (even electron count) or UHF (odd) through the GPU
// engine and writes the reference's output files there.  Optional arguments:
//   --diis   Pulay DIIS instead of the reference damping
//   -v       per-iteration lines on stderr
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "../../../include/unomol.h"

int main(int argc, char** argv) {
    uno_scf_opts o;
    uno_scf_defaults(&o);
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--diis")) o.accel = 2;
        else if (!std::strcmp(argv[i], "-v")) o.verbose = 1;
    }
    uno_scf_result r;
    const int rc = uno_run_deck(".", &o, &r);
    if (rc && rc != UNO_ENOCONV) {
        // the reference prints the message and exits (Util.cpp:6-9)
        std::fprintf(stderr, "%s\n", uno_last_error());
        return EXIT_FAILURE;
    }
    std::fprintf(stderr, "SCF time = %.6f s  (%d iterations, Fock %.6f s, diag %.6f s)\n", r.wall_s, r.iterations,
                 r.fock_s, r.diag_s);
    return EXIT_SUCCESS;
}
