/* Problem description shared by the CPU oracle and the reference harness.
 *
 * Test infrastructure only: the oracle and the reference harness are checkers.
 * Nothing in the shipped library includes this header.
 *
 * The arrays have exactly the layouts lioamber passes to g2g_parameter_init_
 * (g2g/init.cpp:69-219): column-major Fortran arrays, 1-based Nuc, packed
 * lower-triangle density with doubled off-diagonals.
 */
#ifndef XC_INPUT_H
#define XC_INPUT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int natom;               /* QM atoms (max_atoms == natom here)            */
  int M;                   /* basis functions (== ngaussians)               */
  int grid_type;           /* 0: 50, 1: 116, 2: 194 angular points          */
  int iexch;               /* functional id, 1..9 (9 = PBE)                  */
  int normalize;           /* NORM: d xx/yy/zz carry 1/sqrt(3)               */
  const double* r;         /* natom x 3, column-major, ld = natom (bohr)     */
  const unsigned* Iz;      /* natom atomic numbers                           */
  const double* Rm;        /* size parameter by Z (0..54)                     */
  const unsigned* Nr;      /* radial shells by Z for grid type 0             */
  const unsigned* Nr2;     /* radial shells by Z for grid types 1 and 2      */
  const unsigned* Nuc;     /* M, 1-based atom of each function               */
  const unsigned* ncont;   /* M, primitives per function                      */
  const unsigned* nshell;  /* 3: #s, #p functions, #d functions              */
  const double* a;         /* M x 13 column-major exponents                  */
  const double* c;         /* M x 13 column-major coefficients               */
  const double* rho;       /* M(M+1)/2 packed density (off-diagonals x2)     */
  const double* e[3];      /* angular unit vectors, (n x 3) col-major       */
  const double* wang[3];   /* angular weights (sum 4 pi)                     */
  /* g2g_set_options_ knobs (g2g/init.cpp:394-433) */
  double little_cube_size;
  double sphere_radius;
  int assign_all_functions;
  int remove_zero_weights;
  unsigned min_points_per_cube;
  unsigned max_function_exponent;
  /* open shell: rho is the alpha density, rho_b the beta one (same packing) */
  int open;
  const double* rho_b;
} XcInput;

/* One point group as the partition defines it. */
typedef struct {
  int kind;                /* 0 cube, 1 sphere                                */
  int npoints;             /* after zero-weight removal                      */
  int nfuncs;              /* expanded functions s + 3p + 6d                 */
  int s, p, d;             /* shells of each kind selected                   */
  int nnuc;                /* atoms owning selected functions                */
  double wsum;             /* sum of final point weights                     */
  double xsum;             /* sum of point x coordinates (placement check)   */
  unsigned long long fhash;/* FNV-1a over the sorted global function list    */
} XcGroupInfo;

#ifdef __cplusplus
}
#endif
#endif
