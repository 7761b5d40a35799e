/*
 * CPU oracle for the gauged-string-method relaxation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference solver's arithmetic (mepgl/src/nlcg.cu, interpolation.cu,
 * parallel_sum.cu, and the relaxation loop of mepgl.cu).  It exists to check
 * the HIP product path and to serve as the timed CPU baseline.  The product
 * library never links, loads or calls it.
 *
 * Memory layout is the reference's device layout: complex/vector fields are
 * interleaved pairs (real2), index  i + N*j + N*N*n  with i along x, j along y
 * and n the image (io.cu:91-125).  Per-site parameter planes and the
 * neighbour maps are N*N, shared by all images.
 */
#ifndef MEPGL_ORACLE_H
#define MEPGL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Energy-term switches, the runtime form of the NO_NL / NO_KINETIC / NO_EM
 * debug macros of nlcg.cu:38-41. */
enum { OR_TERM_NL = 1, OR_TERM_KIN = 2, OR_TERM_EM = 4, OR_TERM_ALL = 7 };

typedef struct {
    int N;                 /* grid points per side (config.cuh:23)          */
    double dx;             /* lattice spacing (config.cuh:33)               */
    int multicomponent;    /* MULTICOMPONENT (config.cuh:17)                */
    int no_self_field;     /* NO_SELF_FIELD (config.cuh:20, nlcg.cu:33)     */
    int terms;             /* OR_TERM_* mask, OR_TERM_ALL for the solver    */
    int poly_quirks;       /* 1: reproduce nlcg.cu:1162-1165,1251 (Q1, Q2)  */
    const double *a_1, *b_1, *m_xx_1, *m_yy_1;
    const double *a_2, *b_2, *m_xx_2, *m_yy_2;
    const double *h;
    double q_1, q_2, eta, gamma, delta;
    const int *comp_nn;    /* buildNearestNeighbourMap of comp_domain      */
    const int *sc_nn;      /* buildNearestNeighbourMap of sc_domain        */
} or_system;

/* Solver state of one string, named as in mepgl.cu:38-140.  Field arrays are
 * 2*F*N*N doubles.  g_a[0]/g_p1[0]/g_p2[0] is the caller's dev_*_g and
 * index 1 its dev_*_g_old; g_norm2[0] is the caller's g_norm2 and
 * g_norm2[1] its g_norm2_old (per image, F each). */
typedef struct {
    int F;
    double *a, *psi_1, *psi_2;
    double *g_a[2], *g_p1[2], *g_p2[2];
    double *d_a, *d_p1, *d_p2;
    double *g_norm2[2];
    double *g_delta_g, *alpha, *fenergy, *poly; /* poly: 4*F (B,C,D,E) */
} or_state;

void or_build_nn_map(int N, const int *domain, int *nn_map);

void or_free_energy_density(const or_system *s, int F, const double *a,
                            const double *psi_1, const double *psi_2,
                            double *density);
void or_free_energy(const or_system *s, int F, const double *a,
                    const double *psi_1, const double *psi_2, double *fenergy);

void or_gradient(const or_system *s, int F, const double *a,
                 const double *psi_1, const double *psi_2, double *a_g,
                 double *psi_1_g, double *psi_2_g, int maxwell_solver);

void or_compute_jb(const or_system *s, int F, const double *a,
                   const double *psi_1, const double *psi_2, double *b,
                   double *j_1, double *j_2);

void or_poly_expansion(const or_system *s, int F, const double *a,
                       const double *psi_1, const double *psi_2,
                       const double *a_d, const double *psi_1_d,
                       const double *psi_2_d, double *fenergy_density,
                       double *poly_density);

double or_line_search(double A, double B, double C, double D, double E);

void or_nlcg_steps(const or_system *s, or_state *st, int iter_n, int nlcg,
                   int maxwell_solver);

void or_arc_length(int N, int F, int multicomponent, const double *a,
                   const double *psi_1, const double *psi_2, double *s);
void or_arc_length_jb(int N, int F, int multicomponent, const double *b,
                      const double *j_1, const double *j_2, double *s_jb);

void or_redistribute_linear(int N, int F, const double *s, const double *f,
                            int F_new, double *f_new);
int or_redistribute_cubic(int N, int F, const double *s, const double *f,
                          int F_new, double *f_new, int textbook);

void or_gauge_align(int N, int F, int multicomponent, double *psi_1, double *psi_2,
                    double *theta);

void or_thomas(int n, const double *dl, const double *d, const double *du,
               double *rhs);

void or_gauge_align_local(int N, int F, int multicomponent, double dx, double q1,
                          double q2, double *a, double *psi_1, double *psi_2,
                          double *const *rot_1, double *const *rot_2, int nrot,
                          double *chi);

#ifdef __cplusplus
}
#endif
#endif
