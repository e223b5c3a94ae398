/* g2g C ABI of the MI355X XC quadrature engine (libg2g.so).
 *
 * Drop-in for the exchange-correlation library lioamber links with
 * `-lg2g` (lioamber/Makefile.options:48,53).  Every symbol uses the gfortran
 * by-reference convention with a trailing underscore, has the signature of the
 * reference symbol it replaces, and keeps its semantics:
 *   - the arrays passed to g2g_parameter_init_ are referenced, not copied, and
 *     must stay alive until g2g_deinit_ (g2g/init.cpp:115-200);
 *   - g2g_solve_groups_ ASSIGNS *energy and ACCUMULATES (+=) the XC Fock into
 *     the packed fock_vec given at init (g2g/partition.cpp:559-583);
 *   - errors terminate the process with a message, as the reference's
 *     uncaught exceptions and NaN check do (g2g/init.cpp:142,262;
 *     g2g/partition.cpp:584-590).
 * Types: `unsigned` is Fortran INTEGER (4 bytes), `bool` is LOGICAL(1)-sized
 * C++ bool as in the reference.
 */
#ifndef LIO_AMD_G2G_H
#define LIO_AMD_G2G_H

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#endif

/* Replaces g2g_init_ (g2g/init.cpp:38-61): selects the HIP device. */
void g2g_init_(void);

/* Replaces g2g_set_options_ (g2g/init.cpp:409-433). */
void g2g_set_options_(double* free_global_memory, double* little_cube_size,
                      double* sphere_radius, bool* assign_all_functions,
                      bool* energy_all_iterations, bool* remove_zero_weights,
                      unsigned* min_points_per_cube, unsigned* max_function_exponent,
                      unsigned* timer_level, unsigned* verbose);

/* Replaces g2g_parameter_init_ (g2g/init.cpp:69-219). */
void g2g_parameter_init_(const unsigned* norm, const unsigned* natom,
                         const unsigned* max_atoms, const unsigned* ngaussians,
                         double* r, double* Rm, const unsigned* Iz,
                         const unsigned* Nr, const unsigned* Nr2, unsigned* Nuc,
                         const unsigned* M, unsigned* ncont,
                         const unsigned* nshell, double* c, double* a,
                         double* rho_vec, double* fock_vec, double* fockb_vec,
                         double* rhoalpha, double* rhobeta, const unsigned* nco,
                         bool* OPEN, const unsigned* nunp, const unsigned* nopt,
                         const unsigned* Iexch, double* e, double* e2,
                         double* e3, double* wang, double* wang2, double* wang3,
                         bool* use_libxc, const unsigned* ex_functional_id,
                         const unsigned* ec_functional_id);

/* Replaces g2g_reload_atom_positions_ (g2g/init.cpp:282-306): re-reads the
 * positions and rebuilds grid and partition (grid_type 0/1/2 = 50/116/194). */
void g2g_reload_atom_positions_(const unsigned* grid_type);

/* Replaces g2g_new_grid_ (g2g/init.cpp:308-315): rebuilds only when the grid
 * type changed. */
void g2g_new_grid_(const unsigned* grid_type);

/* Replaces g2g_solve_groups_ (g2g/init.cpp:332-391).  computation_type:
 * 0 Fock (energy too when energy_all_iterations), 1 energy only,
 * 2 energy + forces (forces are added to `forces`, natom x 3 with leading
 * dimension max_atoms), 3 forces only.  Open shell (OPEN in
 * g2g_parameter_init_) reads rhoalpha/rhobeta and adds to fock_vec and
 * fockb_vec.  *energy is assigned; fock_vec and forces are accumulated. */
void g2g_solve_groups_(const unsigned* computation_type, double* energy, double* forces);

/* Replaces g2g_deinit_ (g2g/init.cpp:221-236). */
void g2g_deinit_(void);

/* Fortran named timers (g2g/timer.cpp:125-296).  Single timers print when
 * stopped if timer_level is odd; summary timers are kept when timer_level > 1
 * and printed as a tree by g2g_timer_summary_. */
void g2g_timer_start_(const char* name, unsigned len);
void g2g_timer_stop_(const char* name, unsigned len);
void g2g_timer_pause_(const char* name, unsigned len);
void g2g_timer_sum_start_(const char* name, unsigned len);
void g2g_timer_sum_stop_(const char* name, unsigned len);
void g2g_timer_sum_pause_(const char* name, unsigned len);
void g2g_timer_clear_(void);
void g2g_timer_summary_(void);

/* Analytic-integral symbols lioamber references (g2g/analytic_integral/
 * aint_init.cpp:38-361).  This library provides no GPU integrals:
 * aint_query_gpu_level_ reports 0 so lioamber keeps its CPU integrals, as
 * the reference's CPU build does (aint_init.cpp:342-345); the others are
 * inert. */
void aint_query_gpu_level_(int* gpu_level);
void aint_parameter_init_(const unsigned* Md, unsigned* ncontd, const unsigned* nshelld,
                          double* cd, double* ad, unsigned* Nucd, double* af,
                          double* G_inv, double* Hmat_vec, double* str, double* fac,
                          double* rmax, unsigned* atomZ_i, const int* level_of_gpu);
void aint_deinit_(void);
void aint_new_step_(void);
void aint_qmmm_init_(const unsigned* nclatom, double* r_all, double* pc);
void aint_coulomb_init_(void);
void aint_qmmm_forces_(double* qm_forces, double* mm_forces);
void aint_qmmm_fock_(double* Es, double* Ens);
void aint_coulomb_forces_(double* qm_forces);
void aint_coulomb_fock_(double* Es);

#ifdef __cplusplus
}
#endif
#endif
