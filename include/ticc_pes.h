/* C-ABI of a compiled potential-energy surface, loaded by
 * pyticc_amd.pes.native.load_native_pes.
 *
 * Replaces the compiled-PES seam of the reference: the f2py interaction routine
 * called as routine(RR, coordinates, *args) per radial point
 * (src/pyticc/pes/fortran/executor.py:283-288, 299-397) and the monomer
 * potentials of PESWrapper (src/pyticc/pes/adiabatic.py:17-47).
 *
 * Coordinates are column-major, as the reference hands them to Fortran:
 * coords[k + n_coord * g] is coordinate k of grid point g.  Atom-diatom rows
 * are (r, theta); diatom-diatom rows are (r_X, r_Y, theta_X, theta_Y, phi);
 * angles in radians, distances and energies in atomic units.
 *
 * Every entry point returns 0 on success and a nonzero code on failure (the
 * loader raises RuntimeError with that code).  With processes > 1 the loader
 * calls the library from several threads at once, so a library used that way
 * must be re-entrant.
 */
#ifndef TICC_PES_H
#define TICC_PES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Required: V(R, coords[:, g]) for the n_grid points, into out[g]. */
int pes_interaction(double R, const double* coords, int n_coord, int n_grid, double* out);

/* Optional: n_R radial points at once, out row-major (n_R x n_grid). */
int pes_interaction_many(const double* R, int n_R, const double* coords, int n_coord, int n_grid, double* out);

/* Optional monomer potentials: v(r[i]) into out[i] (symbol names are chosen by the caller). */
int pes_monomer(const double* r, int n, double* out);

#ifdef __cplusplus
}
#endif

#endif
