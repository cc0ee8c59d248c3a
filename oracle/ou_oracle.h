/*
 * ou_oracle.h -- FP64 CPU restatement of the Ouranos DG-Sn transport sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in ouranos_amd/.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every routine restates a routine of the reference (Rombur/ouranos, C++
 * deal.II/Trilinos code) in plain C99 on a structured Cartesian mesh:
 *
 *   orc_fe_cell          FECell.cc:10-78           (full Gauss quadrature, like FEValues)
 *   orc_lu_decompose     RadiativeTransfer.cc:1026-1065 (Crout, row partial pivoting)
 *   orc_lu_solve         RadiativeTransfer.cc:1067-1105
 *   orc_ls_octant        LS.cc:14-169,667-682      (level-symmetric octant tables S2..S12)
 *   orc_glc_octant       GLC.cc:14-58              (triangular Gauss-Legendre-Chebyshev)
 *   orc_deploy           RTQuadrature.cc:58-95     (quadrant mirror; 3D adds the mu<0 half)
 *   orc_harmonics        RTQuadrature.cc:97-198    (real spherical harmonics, M2D/D2M)
 *   orc_most_normal      RTQuadrature.cc:200-238
 *   orc_scatter_source   RadiativeTransfer.cc:811-843
 *   orc_sweep            RadiativeTransfer.cc:878-1012 (wavefront order, one cell per task)
 *   orc_apply            RadiativeTransfer.cc:780-809
 *   orc_si / orc_gmres / orc_bicgstab   ouranos.cc:43-209 (driver semantics)
 *
 * Deviations from the shipped reference follow its intended mathematics; each
 * is listed in the header comment of the routine concerned.
 *
 * Layout conventions (shared with the product so vectors can be compared):
 *   cell index  c = i + nx*(j + ny*k)
 *   local dof   l = lx + (p+1)*(ly + (p+1)*lz)     (lexicographic, x fastest)
 *   dof         c*N + l,  N = (p+1)^dim
 *   moments     x[m*n_dofs + dof]                   (moment-major)
 */
#ifndef OU_ORACLE_H
#define OU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_VACUUM = 0, ORC_BC_ISOTROPIC = 1, ORC_BC_MOST_NORMAL = 2 };

typedef struct {
  int dim;            /* 2 or 3 */
  int p;              /* FE degree (1..5) */
  int n[3];           /* cells per axis (n[2]=1 in 2D) */
  double h[3];        /* cell sizes */
  int n_dir;
  const double *omega;   /* n_dir x 3 */
  int n_mom;
  const double *M2D;     /* n_dir x n_mom, row-major */
  const double *D2M;     /* n_mom x n_dir, row-major */
  const int *mom_l;      /* n_mom: Legendre order of each moment */
  int n_mat;
  int L;                 /* scattering expansion order held by sigma_s */
  const double *sigma_t; /* n_mat */
  const double *sigma_s; /* n_mat x (L+1) */
  const int *mat;        /* n_cells */
  const double *q;       /* n_cells: isotropic source intensity (may be NULL) */
  double ws;             /* weight sum */
  int bc[6];             /* per face: ORC_BC_* */
  double inc[6];         /* incoming flux value per face */
  const unsigned char *most_normal; /* 2*dim x n_dir flags (may be NULL) */
} orc_problem;

/* ---- quadrature ---------------------------------------------------------- */
void orc_gauss_legendre(int n, double *x, double *w);
int orc_ls_octant(int sn, double *omega, double *w);
int orc_glc_octant(int sn, double *omega, double *w);
int orc_product_octant(int n_polar, int n_azim, double *omega, double *w);
int orc_deploy(int dim, int n_oct, double *omega, double *w);
int orc_harmonics(int dim, int L, int galerkin, int sn, int n_dir,
                  const double *omega, const double *w, double ws,
                  double *M2D, double *D2M, int *mom_l);
void orc_most_normal(int dim, int n_dir, const double *omega,
                     unsigned char *flags);

/* ---- finite elements ----------------------------------------------------- */
int orc_n_loc(int dim, int p);
void orc_fe_cell(int dim, int p, const double *h, double *M, double *G,
                 double *E, double *U, double *normals);

/* ---- dense kernels ------------------------------------------------------- */
void orc_lu_decompose(int n, double *A, int *pivot);
void orc_lu_solve(int n, const double *A, double *b, double *x,
                  const int *pivot);

/* ---- operator ------------------------------------------------------------ */
long orc_n_dofs(const orc_problem *pb);
void orc_scatter_source(const orc_problem *pb, const double *x, double *S);
void orc_sweep(const orc_problem *pb, const double *S, int with_source,
               double *phi);
void orc_sweep_dirs(const orc_problem *pb, const double *S, int with_source,
                    int d_begin, int d_end, double *phi);
void orc_apply(const orc_problem *pb, const double *x, double *y);
void orc_rhs(const orc_problem *pb, double *rhs);

/* ---- solvers: return iterations used (negative if not converged) -------- */
int orc_si(const orc_problem *pb, double tol, int max_it, double *phi);
int orc_gmres(const orc_problem *pb, const double *b, double *x, int restart,
              double tol, int max_it, double *final_rel_res);
int orc_bicgstab(const orc_problem *pb, const double *b, double *x,
                 double tol, int max_it, double *final_rel_res);

#ifdef __cplusplus
}
#endif
#endif
