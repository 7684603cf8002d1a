"""unomol_amd -- MI355X (gfx950) RHF/UHF SCF engine, a drop-in for the integrals +
Fock build + diagonalisation path of PatNichols/unomol.

The compute path is libunomol_hip.so (hand-written HIP kernels behind a C ABI); this
package only binds it.  See ``api`` for the mirror of the reference's class surface
(Basis, TwoElectronInts, SymmPack, RestrictedHartreeFock, UnRestrictedHartreeFock).
"""
from .api import (Basis, TwoElectronInts, SymmPack, RestrictedHartreeFock,  # noqa: F401
                  UnRestrictedHartreeFock, run_deck, generate_deck, rys_eval, rys_eval_ref2, diis_coefficients, oneint, moment_ints,
                  scf_options, shell_rotation, guess_features, guess_samples, guess_assemble, MLP,
                  set_guess_model, Batch)
