/*
 * mercury_oracle.c -- CPU restatement of the MPCNS/Mercury explicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * file: it is the checker that the HIP path is compared against in tests/, in
 * __graft_entry__.smoke() and in the cpu_baseline leg of bench.py.
 *
 * Every function restates one reference function and cites it (paths are
 * relative to the reference repository root).  The reference's Mercury TUs all
 * include <petscksp.h> (Z4_Mercury/include/MercurySolver.h:14) and the image has
 * no PETSc, so the reference cannot be compiled here; this restatement is the
 * oracle and is pinned by the closed-form / invariant / known-answer tests under
 * tests/ (see the header of each function for what pins it).
 *
 * Compile with -ffp-contract=off: the reference is built without FMA
 * contraction on x86-64, and the device FP64 path is built the same way, so the
 * two can be compared bit for bit wherever they perform the same operations.
 *
 * Layout: per block, every field component is one array over the box
 * [-g, m+g] on each axis (m cells), k fastest:  idx = ((i+g)*NJ+(j+g))*NK+(k+g).
 * Node coordinates live on a larger box [-g-1, m+g+1] (ghost depth g+1), like
 * the reference grid arrays (src/1_grid/1_MPCNS_Grid.cpp:92-101).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------- */
/* physics constants, same layout as MercuryPhysics in include/mercury_hip.h   */
/* ------------------------------------------------------------------------- */
typedef struct {
    double gamma;
    double M_H, M_Na, M_ref;
    double state_coeff_H, state_coeff_Na;
    double T_ref;
    double inver_MA2;
    double hall_coef;
    double ne_floor;
    double c_ind, c_hall;
    double a1_Na, a4, a5, a6;
    double Tn0;
    double nu_lit, nu_shadow;
    double cfl;
    double taper_rmin, taper_rmax;
    double arti_eta_max, J_start, J_on;
    double local_enabled, local_eta_max, local_center[3], local_r_decay, local_r_cutoff;
    double ambipolar_enabled, pressure_coef;
    double qinf[5], qinfs[5], q_pv_inf[5], B_imf[3];
    double wall_rho_floor_H, wall_rho_floor_Na, wall_p_floor;
    /* implicit resistive correction (MercurySolver.cpp:173-184, MercurySolver_Tools.cpp:44) */
    double resist_enabled, inver_Rem, resist_r_inner, resist_r_outer, resist_width;
    double resist_rtol, resist_atol, resist_max_it;
    double resist_profile; /* 0: the reference's window, 1: exponential (no reference counterpart) */
    /* ion species: 2 (H+, Na+), the reference's two fluids (Z4_Mercury/include/0_MercuryFieldCatalog.h:77-82),
     * or 3 with He++ (no reference counterpart): its mass, state coefficient, charge
     * number, far-field state and wall density floor */
    double n_species, M_He, state_coeff_He, Z_He, qinf_He[5], wall_rho_floor_He;
} OrcPhys;

/* face codes; BC_INTERFACE: a 1:1 inner connection to another (or the same)
 * block, given by the neighbour and a translation of node indices */
enum { BC_PERIODIC = 0, BC_FARFIELD = 1, BC_WALL = 2, BC_INTERFACE = 3 };

/* stagger locations */
enum { L_CELL = 0, L_FX, L_FY, L_FZ, L_EX, L_EY, L_EZ, L_NODE };
static const int DELTA[8][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0},
                                {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};

/* field slots ------------------------------------------------------------- */
enum {
    /* cell */
    S_UH = 0, S_UN = S_UH + 5, S_PVH = S_UN + 5, S_PVN = S_PVH + 5,
    S_UP = S_PVN + 5, S_BC = S_UP + 3, S_BI = S_BC + 3, S_JC = S_BI + 3,
    S_RH = S_JC + 3, S_RN = S_RH + 5, S_EHC = S_RN + 5, S_HB = S_EHC + 3,
    S_JAC = S_HB + 1, S_WB = S_JAC + 1, S_WJ = S_WB + 18, S_PHOTO = S_WJ + 36,
    S_CC = S_PHOTO + 1, S_DIVB = S_CC + 3, S_U0H = S_DIVB + 1, S_U0N = S_U0H + 5,
    /* faces (3 axes each) */
    S_B = S_U0N + 5, S_BADD = S_B + 3, S_RHSB = S_BADD + 3, S_JD = S_RHSB + 3,
    S_AREA = S_JD + 9, S_BETA = S_AREA + 3, S_EF = S_BETA + 3, S_F = S_EF + 9,
    S_B0 = S_F + 15,
    /* edges */
    S_E = S_B0 + 3, S_EH = S_E + 3, S_J = S_EH + 3, S_DR = S_J + 3, S_DL = S_DR + 9,
    S_ALPHA = S_DL + 3, S_TAPER = S_ALPHA + 3, S_LETA = S_TAPER + 3,
    /* generic DEC forms: node 0-form, edge 1-forms, face 2-forms, cell 3-form */
    S_NODEF = S_LETA + 3, S_EDGEF = S_NODEF + 1, S_FACEF = S_EDGEF + 3, S_CELLS = S_FACEF + 3,
    /* implicit resistive correction: Eres, dJ, eta on the unknown edges, the
     * solution kept as the next initial guess (edges); dB, RHS_Bres (faces) */
    S_ERES = S_CELLS + 1, S_DJ = S_ERES + 3, S_ETAR = S_DJ + 3, S_XRES = S_ETAR + 3,
    S_DB = S_XRES + 3, S_RHSBR = S_DB + 3,
    /* the third ion species (He++): state, PV, RHS and step-start copy */
    S_UHE = S_RHSBR + 3, S_PVHE = S_UHE + 5, S_RHE = S_PVHE + 5, S_U0HE = S_RHE + 5,
    S_NSLOT = S_U0HE + 5
};

/* species order H+, Na+, He++: the reference's two species first, so every sum
 * over species adds the He++ term last and two-species runs are unchanged */
#define NSMAX 3
static const int SP_U[NSMAX] = {S_UH, S_UN, S_UHE}, SP_PV[NSMAX] = {S_PVH, S_PVN, S_PVHE}, SP_R[NSMAX] = {S_RH, S_RN, S_RHE},
                 SP_U0[NSMAX] = {S_U0H, S_U0N, S_U0HE};
static int n_species(const OrcPhys *P) { return P->n_species >= 3.0 ? 3 : 2; }
static double species_mass(const OrcPhys *P, int s) { return s == 0 ? P->M_H : (s == 1 ? P->M_Na : P->M_He); }
static double species_coeff(const OrcPhys *P, int s) { return s == 0 ? P->state_coeff_H : (s == 1 ? P->state_coeff_Na : P->state_coeff_He); }

typedef struct {
    int m[3], g;
    int N[3];
    long V;
    int bc[6];
    int nbr[6];    /* BC_INTERFACE: neighbour block of the face */
    int off[6][3]; /* BC_INTERFACE: node (i,j,k) -> neighbour node (i,j,k) + off */
    int Nn[3];
    long Vn;
    double *x, *y, *z;
    double *f[S_NSLOT];
} OBlock;

typedef struct {
    int nb;
    OBlock *b;
    OrcPhys P;
    double dt;
    double time;
    int step;
    int integrator; /* 0 Euler (reference), 1 SSP-RK2 */
    /* equivalence classes of shared DOFs per location (faces, edges, nodes):
     * members in ascending EntityKey order (block, i, j, k), the first is the
     * owner; cls_b / cls_q: block and box index of each member */
    int ncls[8];
    long *cls_start[8];
    int *cls_b[8];
    long *cls_q[8];
    /* each member's axis (the field of its location within a face or edge
     * triplet) and orientation sign; with patch connections a class may hold
     * members of different axes, and all classes of a dimension are listed
     * under its first location (L_FX, L_EX) */
    int *cls_ax[8];
    signed char *cls_sg[8];
    /* patch connections (0 when the faces carry per-face connections): 18 ints
     * per patch (block, dir, node box lo[3] hi[3], neighbour block, perm[3],
     * sign[3], offset[3]) and the halo regions of each level built from them */
    int npatch;
    int *patch;
    int rule; /* 0: the reference's node-index map, 1: DOF positions matched */
    int nreg[4];
    int *reg[4];
    /* during a patch sync: per field axis and block, the box points the halo
     * regions wrote (rule 1 completes the others) */
    char *wm[3][64];
    /* singular edges (patch contexts): entries, their cells and faces */
    struct OSingRegT *singp;
} Orc;

#define IX(b, i, j, k) ((((long)(i) + (b)->g) * (b)->N[1] + ((j) + (b)->g)) * (b)->N[2] + ((k) + (b)->g))
static void singular_current(Orc *o, int sB, int sJ);
#define NX(b, i, j, k) ((((long)(i) + (b)->g + 1) * (b)->Nn[1] + ((j) + (b)->g + 1)) * (b)->Nn[2] + ((k) + (b)->g + 1))
#define F(b, s, i, j, k) ((b)->f[(s)][IX(b, i, j, k)])
#define XN(b, i, j, k) ((b)->x[NX(b, i, j, k)])
#define YN(b, i, j, k) ((b)->y[NX(b, i, j, k)])
#define ZN(b, i, j, k) ((b)->z[NX(b, i, j, k)])

/* Threading: every per-point loop marked PAR2 / PAR1 writes each output element
 * from inputs held in other slots, so its iterations are independent and a
 * threaded run gives the same bits as a serial one.  PAR2 splits the two outer
 * levels (i, j) of a perfectly nested i/j/k loop, PAR1 a flat loop over the
 * box.  Halo / boundary fills, alias means and the resistive CG (whose dot
 * products would change with the summation order) stay serial.  The thread
 * count is OpenMP's (OMP_NUM_THREADS). */
#define PAR2 _Pragma("omp parallel for collapse(2) schedule(static)")
#define PAR1 _Pragma("omp parallel for schedule(static)")

static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* owned range of a location: [0, m+1-delta)  (include/3_field/FieldBlock.h:34-35) */
static void owned_hi(const OBlock *b, int loc, int hi[3]) {
    for (int a = 0; a < 3; ++a) hi[a] = b->m[a] + 1 - DELTA[loc][a];
}
/* allocated range with ghost depth G (src/0_basic/LayoutTraits.cpp:278-294) */
static void alloc_box(const OBlock *b, int loc, int G, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = -G;
        hi[a] = b->m[a] + 1 - DELTA[loc][a] + G;
    }
}

/* ------------------------------------------------------------------------- */
/* geometry                                                                   */
/* ------------------------------------------------------------------------- */

/* Block::Add_ghostmesh, 3-D branch (src/1_grid/1_MPCNS_Grid.cpp:591-818).
 * Ghost nodes are mirror-extrapolated one layer at a time; periodic interfaces
 * do not overwrite them (src/1_grid/1_MPCNS_Grid.cpp:1268-1269). */
static void add_ghostmesh(OBlock *b) {
    const int mx = b->m[0], my = b->m[1], mz = b->m[2];
    double *X[3] = {b->x, b->y, b->z};
    for (int gh = 1; gh <= b->g + 1; ++gh) {
        const int gimin = -gh, gimax = mx + gh, gjmin = -gh, gjmax = my + gh, gkmin = -gh, gkmax = mz + gh;
        for (int c = 0; c < 3; ++c) {
            double *q = X[c];
#define Q(i, j, k) q[NX(b, i, j, k)]
            for (int j = gjmin + 1; j < gjmax; j++)
                for (int k = gkmin + 1; k < gkmax; k++) {
                    Q(gimin, j, k) = 2 * Q(0, j, k) - Q(-gimin, j, k);
                    Q(gimax, j, k) = 2 * Q(mx, j, k) - Q(2 * mx - gimax, j, k);
                }
            for (int i = gimin + 1; i < gimax; i++)
                for (int k = gkmin + 1; k < gkmax; k++) {
                    Q(i, gjmin, k) = 2 * Q(i, 0, k) - Q(i, -gjmin, k);
                    Q(i, gjmax, k) = 2 * Q(i, my, k) - Q(i, 2 * my - gjmax, k);
                }
            for (int i = gimin + 1; i < gimax; i++)
                for (int j = gjmin + 1; j < gjmax; j++) {
                    Q(i, j, gkmin) = 2 * Q(i, j, 0) - Q(i, j, -gkmin);
                    Q(i, j, gkmax) = 2 * Q(i, j, mz) - Q(i, j, 2 * mz - gkmax);
                }
            for (int k = gkmin + 1; k < gkmax; k++) {
                Q(gimin, gjmin, k) = Q(gimin + 1, gjmin, k) + Q(gimin, gjmin + 1, k) - Q(gimin + 1, gjmin + 1, k);
                Q(gimax, gjmin, k) = Q(gimax - 1, gjmin, k) + Q(gimax, gjmin + 1, k) - Q(gimax - 1, gjmin + 1, k);
                Q(gimin, gjmax, k) = Q(gimin + 1, gjmax, k) + Q(gimin, gjmax - 1, k) - Q(gimin + 1, gjmax - 1, k);
                Q(gimax, gjmax, k) = Q(gimax - 1, gjmax, k) + Q(gimax, gjmax - 1, k) - Q(gimax - 1, gjmax - 1, k);
            }
            for (int j = gjmin + 1; j < gjmax; j++) {
                Q(gimin, j, gkmin) = Q(gimin + 1, j, gkmin) + Q(gimin, j, gkmin + 1) - Q(gimin + 1, j, gkmin + 1);
                Q(gimax, j, gkmin) = Q(gimax - 1, j, gkmin) + Q(gimax, j, gkmin + 1) - Q(gimax - 1, j, gkmin + 1);
                Q(gimin, j, gkmax) = Q(gimin + 1, j, gkmax) + Q(gimin, j, gkmax - 1) - Q(gimin + 1, j, gkmax - 1);
                Q(gimax, j, gkmax) = Q(gimax - 1, j, gkmax) + Q(gimax, j, gkmax - 1) - Q(gimax - 1, j, gkmax - 1);
            }
            for (int i = gimin; i < gimax + 1; i++) {
                Q(i, gjmin, gkmin) = Q(i, gjmin + 1, gkmin) + Q(i, gjmin, gkmin + 1) - Q(i, gjmin + 1, gkmin + 1);
                Q(i, gjmax, gkmin) = Q(i, gjmax - 1, gkmin) + Q(i, gjmax, gkmin + 1) - Q(i, gjmax - 1, gkmin + 1);
                Q(i, gjmin, gkmax) = Q(i, gjmin + 1, gkmax) + Q(i, gjmin, gkmax - 1) - Q(i, gjmin + 1, gkmax - 1);
                Q(i, gjmax, gkmax) = Q(i, gjmax - 1, gkmax) + Q(i, gjmax, gkmax - 1) - Q(i, gjmax - 1, gkmax - 1);
            }
#undef Q
        }
    }
}

static inline void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm3(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void node(const OBlock *b, int i, int j, int k, double p[3]) {
    p[0] = XN(b, i, j, k);
    p[1] = YN(b, i, j, k);
    p[2] = ZN(b, i, j, k);
}

/* Block::calc_Dual_Grids (src/1_grid/1_MPCNS_Grid_GCL.cpp:6-24): centre of
 * cell (i,j,k) = mean of its 8 nodes, stored at the cell's own index. */
static void dual_grid(OBlock *b) {
    for (int i = -b->g; i < b->m[0] + b->g; ++i)
        for (int j = -b->g; j < b->m[1] + b->g; ++j)
            for (int k = -b->g; k < b->m[2] + b->g; ++k) {
                double *q[3] = {b->x, b->y, b->z};
                for (int c = 0; c < 3; ++c) {
                    const double *Q = q[c];
#define QQ(a1, a2, a3) Q[NX(b, a1, a2, a3)]
                    F(b, S_CC + c, i, j, k) = 0.125 * (QQ(i + 1, j + 1, k + 1) + QQ(i, j + 1, k + 1) + QQ(i + 1, j, k + 1) + QQ(i, j, k + 1) +
                                                       QQ(i + 1, j + 1, k) + QQ(i, j + 1, k) + QQ(i + 1, j, k) + QQ(i, j, k));
#undef QQ
                }
            }
}

static void cellc(const OBlock *b, int i, int j, int k, double p[3]) {
    p[0] = F(b, S_CC + 0, i, j, k);
    p[1] = F(b, S_CC + 1, i, j, k);
    p[2] = F(b, S_CC + 2, i, j, k);
}

/* quad area vector, 0.5*((p10-p00)x(p01-p00) + (p01-p11)x(p10-p11))
 * (src/7_metric/Metric.cpp:109-141 and :380-389) */
static void quad_area(const double p00[3], const double p10[3], const double p01[3], const double p11[3], double A[3]) {
    double r1[3], r2[3], r3[3], r4[3], c1[3], c2[3];
    for (int c = 0; c < 3; ++c) {
        r1[c] = p10[c] - p00[c];
        r2[c] = p01[c] - p00[c];
        r3[c] = p01[c] - p11[c];
        r4[c] = p10[c] - p11[c];
    }
    cross3(r1, r2, c1);
    cross3(r3, r4, c2);
    for (int c = 0; c < 3; ++c) A[c] = 0.5 * (c1[c] + c2[c]);
}

/* METRIC::compute_metric_fields (src/7_metric/Metric.cpp:78-1414): JD (face
 * area vectors, ghost g), Jac (ghost g-1), Area / face Hodge beta (ghost g-1),
 * dl / edge inverse Hodge alpha with the dual-collapse cap (ghost g-1), dr
 * (ghost g-1). */
static void metric(OBlock *b) {
    const int g = b->g;
    int lo[3], hi[3];
    /* JDxi / JDet / JDze (Metric.cpp:109-211) */
    for (int a = 0; a < 3; ++a) {
        alloc_box(b, L_FX + a, g, lo, hi);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    double o[3], pa[3], pb[3], pab[3];
                    node(b, i, j, k, o);
                    if (a == 0) {
                        node(b, i, j + 1, k, pa); node(b, i, j, k + 1, pb); node(b, i, j + 1, k + 1, pab);
                    } else if (a == 1) {
                        node(b, i, j, k + 1, pa); node(b, i + 1, j, k, pb); node(b, i + 1, j, k + 1, pab);
                    } else {
                        node(b, i + 1, j, k, pa); node(b, i, j + 1, k, pb); node(b, i + 1, j + 1, k, pab);
                    }
                    double rx[3], ry[3], rx_[3], ry_[3], c1[3], c2[3];
                    for (int c = 0; c < 3; ++c) {
                        rx[c] = pa[c] - o[c];
                        ry[c] = pb[c] - o[c];
                        rx_[c] = pb[c] - pab[c];
                        ry_[c] = pa[c] - pab[c];
                    }
                    cross3(rx, ry, c1);
                    cross3(rx_, ry_, c2);
                    for (int c = 0; c < 3; ++c) F(b, S_JD + 3 * a + c, i, j, k) = 0.5 * (c1[c] + c2[c]);
                }
    }
    /* Jac = 1/3 sum_f A_f.(x_f - x_c) (Metric.cpp:214-326) */
    alloc_box(b, L_CELL, g - 1, lo, hi);
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                double xc[3];
                cellc(b, i, j, k, xc);
                double V = 0.0;
                /* faces in the reference order: xi-, xi+, eta-, eta+, zeta-, zeta+ */
                for (int a = 0; a < 3; ++a)
                    for (int s = 0; s < 2; ++s) {
                        int fi = i + (a == 0 ? s : 0), fj = j + (a == 1 ? s : 0), fk = k + (a == 2 ? s : 0);
                        double Af[3];
                        for (int c = 0; c < 3; ++c) {
                            double v = F(b, S_JD + 3 * a + c, fi, fj, fk);
                            Af[c] = s ? v : -v;
                        }
                        /* four corner nodes of the face, averaged in the reference's order */
                        double n0[3], n1[3], n2[3], n3[3], xf[3];
                        if (a == 0) {
                            node(b, fi, j, k, n0); node(b, fi, j + 1, k, n1); node(b, fi, j, k + 1, n2); node(b, fi, j + 1, k + 1, n3);
                        } else if (a == 1) {
                            node(b, i, fj, k, n0); node(b, i + 1, fj, k, n1); node(b, i, fj, k + 1, n2); node(b, i + 1, fj, k + 1, n3);
                        } else {
                            node(b, i, j, fk, n0); node(b, i + 1, j, fk, n1); node(b, i, j + 1, fk, n2); node(b, i + 1, j + 1, fk, n3);
                        }
                        for (int c = 0; c < 3; ++c) xf[c] = 0.25 * (n0[c] + n1[c] + n2[c] + n3[c]);
                        double d[3] = {xf[0] - xc[0], xf[1] - xc[1], xf[2] - xc[2]};
                        V += dot3(Af, d);
                    }
                F(b, S_JAC, i, j, k) = V / 3.0;
            }
    /* Area, face Hodge beta = |l*|/|S| (Metric.cpp:397-489) */
    const double eps = 1e-25;
    for (int a = 0; a < 3; ++a) {
        alloc_box(b, L_FX + a, g - 1, lo, hi);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    double S[3] = {F(b, S_JD + 3 * a, i, j, k), F(b, S_JD + 3 * a + 1, i, j, k), F(b, S_JD + 3 * a + 2, i, j, k)};
                    double Smag = norm3(S);
                    F(b, S_AREA + a, i, j, k) = Smag;
                    double cL[3], cR[3], d[3];
                    cellc(b, i - (a == 0), j - (a == 1), k - (a == 2), cL);
                    cellc(b, i, j, k, cR);
                    for (int c = 0; c < 3; ++c) d[c] = cR[c] - cL[c];
                    double Lstar = norm3(d);
                    F(b, S_BETA + a, i, j, k) = (Smag < eps) ? 0.0 : Lstar / Smag;
                }
    }
    /* dl, capped inverse Hodge alpha = |e|/|S*| (Metric.cpp:511-680) */
    const double RATIO = 1e-3, EABS = 1e-12;
    for (int a = 0; a < 3; ++a) {
        alloc_box(b, L_EX + a, g - 1, lo, hi);
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    double r0[3], r1[3], d[3];
                    node(b, i, j, k, r0);
                    node(b, i + (a == 0), j + (a == 1), k + (a == 2), r1);
                    for (int c = 0; c < 3; ++c) d[c] = r1[c] - r0[c];
                    double L = norm3(d);
                    F(b, S_DL + a, i, j, k) = L;
                    /* dual quad: the 4 cells around the edge */
                    double p00[3], p10[3], p01[3], p11[3];
                    if (a == 0) {
                        cellc(b, i, j - 1, k - 1, p00); cellc(b, i, j, k - 1, p10); cellc(b, i, j - 1, k, p01); cellc(b, i, j, k, p11);
                    } else if (a == 1) {
                        cellc(b, i - 1, j, k - 1, p00); cellc(b, i, j, k - 1, p10); cellc(b, i - 1, j, k, p01); cellc(b, i, j, k, p11);
                    } else {
                        cellc(b, i - 1, j - 1, k, p00); cellc(b, i, j - 1, k, p10); cellc(b, i - 1, j, k, p01); cellc(b, i, j, k, p11);
                    }
                    double Av[3];
                    quad_area(p00, p10, p01, p11, Av);
                    double Amag = norm3(Av);
                    double raw = (Amag < eps) ? 0.0 : (L / Amag);
                    double t[3];
#define DIST(P, Qp) (t[0] = Qp[0] - P[0], t[1] = Qp[1] - P[1], t[2] = Qp[2] - P[2], norm3(t))
                    double d00_10 = DIST(p00, p10), d00_01 = DIST(p00, p01), d10_11 = DIST(p10, p11), d01_11 = DIST(p01, p11);
                    double d00_11 = DIST(p00, p11), d10_01 = DIST(p10, p01);
#undef DIST
                    double href = dmax(dmax(dmax(d00_10, d00_01), dmax(d10_11, d01_11)), dmax(dmax(d00_11, d10_01), dmax(L, eps)));
                    double side_min = dmin(dmin(d00_10, d00_01), dmin(d10_11, d01_11));
                    double diag_min = dmin(d00_11, d10_01);
                    int touch = (L < EABS) || (side_min < RATIO * href) || (diag_min < RATIO * href) || (Amag < RATIO * href * href);
                    F(b, S_ALPHA + a, i, j, k) = touch ? 0.0 : raw;
                    for (int c = 0; c < 3; ++c) F(b, S_DR + 3 * a + c, i, j, k) = d[c];
                }
    }
}

/* 3x3 symmetric Jacobi eigen-decomposition with 18 sweeps
 * (Z4_Mercury/src/MercurySolver_ReconstructionSetup.cpp:34-61) */
static void eigen_sym3(double A[3][3], double V[3][3], double d[3]) {
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) V[a][c] = (a == c) ? 1.0 : 0.0;
    for (int it = 0; it < 18; ++it) {
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = A[p][q];
                if (fabs(apq) <= 1.0e-16 * (fabs(A[p][p]) + fabs(A[q][q]) + 1.0)) continue;
                const double phi = 0.5 * atan2(2.0 * apq, A[q][q] - A[p][p]);
                const double c = cos(phi), s = sin(phi);
                const double app = A[p][p], aqq = A[q][q];
                for (int k = 0; k < 3; ++k)
                    if (k != p && k != q) {
                        const double akp = A[k][p], akq = A[k][q];
                        A[k][p] = A[p][k] = c * akp - s * akq;
                        A[k][q] = A[q][k] = s * akp + c * akq;
                    }
                A[p][p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
                A[q][q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
                A[p][q] = A[q][p] = 0.0;
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    d[0] = A[0][0];
    d[1] = A[1][1];
    d[2] = A[2][2];
}

/* SetupCellReconstructionWeights_ (Z4_Mercury/src/MercurySolver_ReconstructionSetup.cpp:63-191):
 * W_b (18 per owned cell) and W_j (36 per owned cell). */
static void recon_weights(OBlock *b) {
    const double eps = 1.0e-30, eig_tol = 1.0e-10;
    for (int i = 0; i < b->m[0]; ++i)
        for (int j = 0; j < b->m[1]; ++j)
            for (int k = 0; k < b->m[2]; ++k) {
                double s[6][3];
                for (int a = 0; a < 3; ++a) {
                    s[0][a] = -F(b, S_JD + 0 + a, i, j, k);
                    s[1][a] = F(b, S_JD + 0 + a, i + 1, j, k);
                    s[2][a] = -F(b, S_JD + 3 + a, i, j, k);
                    s[3][a] = F(b, S_JD + 3 + a, i, j + 1, k);
                    s[4][a] = -F(b, S_JD + 6 + a, i, j, k);
                    s[5][a] = F(b, S_JD + 6 + a, i, j, k + 1);
                }
                double X[8][3];
                int cnt = 0;
                /* node order of avg8 in the reference: (i,j,k),(i+1,j,k),(i,j+1,k),(i+1,j+1,k),(i,j,k+1),... */
                for (int dk = 0; dk < 2; ++dk)
                    for (int dj = 0; dj < 2; ++dj)
                        for (int di = 0; di < 2; ++di) node(b, i + di, j + dj, k + dk, X[cnt++]);
                double xc[3];
                for (int c = 0; c < 3; ++c)
                    xc[c] = 0.125 * (X[0][c] + X[1][c] + X[2][c] + X[3][c] + X[4][c] + X[5][c] + X[6][c] + X[7][c]);
                /* face centres: setfc(n, ...) four nodes each */
                static const int FN[6][4][3] = {
                    {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}}, {{1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
                    {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1}}, {{0, 1, 0}, {1, 1, 0}, {0, 1, 1}, {1, 1, 1}},
                    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}}, {{0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}}};
                double fc[6][3];
                for (int n = 0; n < 6; ++n) {
                    double P[4][3];
                    for (int q = 0; q < 4; ++q) node(b, i + FN[n][q][0], j + FN[n][q][1], k + FN[n][q][2], P[q]);
                    for (int c = 0; c < 3; ++c) fc[n][c] = 0.25 * (P[0][c] + P[1][c] + P[2][c] + P[3][c]);
                }
                double smag[6], mu[6], sref = 0.0;
                for (int n = 0; n < 6; ++n) {
                    smag[n] = sqrt(s[n][0] * s[n][0] + s[n][1] * s[n][1] + s[n][2] * s[n][2]);
                    sref = dmax(sref, smag[n]);
                    const double dx = fc[n][0] - xc[0], dy = fc[n][1] - xc[1], dz = fc[n][2] - xc[2];
                    mu[n] = 1.0 / dmax(sqrt(dx * dx + dy * dy + dz * dz), 1.0e-14);
                }
                double M[3][3] = {{0}};
                int active[6];
                for (int n = 0; n < 6; ++n) {
                    active[n] = smag[n] > 1.0e-10 * sref;
                    if (!active[n]) continue;
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) M[a][c] += mu[n] * s[n][a] * s[n][c];
                }
                double A[3][3], V[3][3], lam[3];
                memcpy(A, M, sizeof(A));
                eigen_sym3(A, V, lam);
                const double lmax = dmax(lam[0], dmax(lam[1], lam[2]));
                for (int m = 0; m < 18; ++m) F(b, S_WB + m, i, j, k) = 0.0;
                if (!(lmax <= eps)) {
                    double Mi[3][3] = {{0}};
                    for (int q = 0; q < 3; ++q) {
                        const double inv = (lam[q] > eig_tol * lmax) ? 1.0 / lam[q] : 0.0;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c) Mi[a][c] += V[a][q] * inv * V[c][q];
                    }
                    for (int n = 0; n < 6; ++n)
                        if (active[n])
                            for (int a = 0; a < 3; ++a)
                                F(b, S_WB + 6 * a + n, i, j, k) = mu[n] * (Mi[a][0] * s[n][0] + Mi[a][1] * s[n][1] + Mi[a][2] * s[n][2]);
                }
                /* --- W_j from the 12 cell edges --- */
                static const int ED[12][6] = {
                    {0, 0, 0, 1, 0, 0}, {0, 1, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 1}, {0, 1, 1, 1, 1, 1},
                    {0, 0, 0, 0, 1, 0}, {1, 0, 0, 1, 1, 0}, {0, 0, 1, 0, 1, 1}, {1, 0, 1, 1, 1, 1},
                    {0, 0, 0, 0, 0, 1}, {1, 0, 0, 1, 0, 1}, {0, 1, 0, 0, 1, 1}, {1, 1, 0, 1, 1, 1}};
                double l[12][3];
                for (int n = 0; n < 12; ++n) {
                    double p0[3], p1[3];
                    node(b, i + ED[n][0], j + ED[n][1], k + ED[n][2], p0);
                    node(b, i + ED[n][3], j + ED[n][4], k + ED[n][5], p1);
                    for (int c = 0; c < 3; ++c) l[n][c] = p1[c] - p0[c];
                }
                double Mj[3][3] = {{0}};
                double lmaxl = 0.0;
                for (int n = 0; n < 12; ++n) lmaxl = dmax(lmaxl, sqrt(l[n][0] * l[n][0] + l[n][1] * l[n][1] + l[n][2] * l[n][2]));
                for (int m = 0; m < 36; ++m) F(b, S_WJ + m, i, j, k) = 0.0;
                if (lmaxl <= eps) continue;
                int act[12];
                for (int n = 0; n < 12; ++n) {
                    const double ln = sqrt(l[n][0] * l[n][0] + l[n][1] * l[n][1] + l[n][2] * l[n][2]);
                    act[n] = ln > 1.0e-10 * lmaxl;
                    if (!act[n]) continue;
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) Mj[a][c] += l[n][a] * l[n][c];
                }
                memcpy(A, Mj, sizeof(A));
                eigen_sym3(A, V, lam);
                const double lme = dmax(lam[0], dmax(lam[1], lam[2]));
                if (lme <= eps) continue;
                double Mi[3][3] = {{0}};
                for (int q = 0; q < 3; ++q) {
                    const double inv = (lam[q] > eig_tol * lme) ? 1.0 / lam[q] : 0.0;
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) Mi[a][c] += V[a][q] * inv * V[c][q];
                }
                for (int n = 0; n < 12; ++n)
                    if (act[n])
                        for (int a = 0; a < 3; ++a)
                            F(b, S_WJ + 12 * a + n, i, j, k) = Mi[a][0] * l[n][0] + Mi[a][1] * l[n][1] + Mi[a][2] * l[n][2];
            }
}

/* HallRadialTaper_/HallRadialTaperEdge_ (MercurySolver_MagExplicit_Hall.cpp:6-25)
 * and the cached local artificial eta (MercurySolver_MagExplicit_Resis.cpp:87-144),
 * both per owned edge. */
static void edge_statics(OBlock *b, const OrcPhys *P) {
    int hi[3];
    for (int a = 0; a < 3; ++a) {
        owned_hi(b, L_EX + a, hi);
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    int ip = i + (a == 0), jp = j + (a == 1), kp = k + (a == 2);
                    double xm = 0.5 * (XN(b, i, j, k) + XN(b, ip, jp, kp));
                    double ym = 0.5 * (YN(b, i, j, k) + YN(b, ip, jp, kp));
                    double zm = 0.5 * (ZN(b, i, j, k) + ZN(b, ip, jp, kp));
                    double r = sqrt(xm * xm + ym * ym + zm * zm);
                    double width = P->taper_rmax - P->taper_rmin, tp;
                    if (width <= 0.0)
                        tp = (r >= P->taper_rmin) ? 1.0 : 0.0;
                    else {
                        double mid = 0.5 * (P->taper_rmin + P->taper_rmax);
                        tp = 0.5 * (1.0 + tanh(2.0 * (r - mid) / width));
                    }
                    F(b, S_TAPER + a, i, j, k) = tp;
                    double eta = 0.0;
                    if (P->local_enabled != 0.0 && P->local_eta_max > 0.0 && P->local_r_decay > 0.0) {
                        double dx = xm - P->local_center[0], dy = ym - P->local_center[1], dz = zm - P->local_center[2];
                        double rr = sqrt(dx * dx + dy * dy + dz * dz);
                        if (P->local_r_cutoff > 0.0 && rr >= P->local_r_cutoff)
                            eta = 0.0;
                        else
                            eta = P->local_eta_max * exp(-rr / P->local_r_decay);
                    }
                    F(b, S_LETA + a, i, j, k) = eta;
                }
    }
}

/* ------------------------------------------------------------------------- */
/* synchronisation: physical handlers, corner default, periodic halo, owner   */
/* ------------------------------------------------------------------------- */

/* which slots belong to which field, with location and component count */
typedef struct {
    int slot, ncomp, loc;
} FDesc;

/* boundary-adjacent one-layer inner slab of a full block face
 * (LAYOUT::boundary_inner_slab_one_layer_from_cells, LayoutTraits.cpp:434-465) */
static void inner_slab(const OBlock *b, int loc, int dir, int lo[3], int hi[3]) {
    int ax = abs(dir) - 1, ohi[3];
    owned_hi(b, loc, ohi);
    for (int a = 0; a < 3; ++a) {
        lo[a] = 0;
        hi[a] = ohi[a];
    }
    if (dir < 0) {
        lo[ax] = 0;
        hi[ax] = 1;
    } else {
        lo[ax] = ohi[ax] - 1;
        hi[ax] = ohi[ax];
    }
}
/* LAYOUT::ghost_slab_from_inner (LayoutTraits.cpp:467-485) */
static void ghost_slab(const int ilo[3], const int ihi[3], int dir, int G, int lo[3], int hi[3]) {
    int ax = abs(dir) - 1;
    for (int a = 0; a < 3; ++a) {
        lo[a] = ilo[a];
        hi[a] = ihi[a];
    }
    if (dir < 0) {
        hi[ax] = ilo[ax];
        lo[ax] = ilo[ax] - G;
    } else {
        lo[ax] = ihi[ax];
        hi[ax] = ihi[ax] + G;
    }
}

/* BoundaryCore::DefaultPhysicalCopy (src/5_boundary/Boundary_Physical.cpp:288-315) */
static void default_copy(OBlock *b, const FDesc *fd, int dir) {
    int ilo[3], ihi[3], lo[3], hi[3];
    inner_slab(b, fd->loc, dir, ilo, ihi);
    ghost_slab(ilo, ihi, dir, b->g, lo, hi);
    int ax = abs(dir) - 1;
    int ref = (dir < 0) ? ilo[ax] : ihi[ax] - 1;
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                int ii = ax == 0 ? ref : i, jj = ax == 1 ? ref : j, kk = ax == 2 ? ref : k;
                for (int m = 0; m < fd->ncomp; ++m) F(b, fd->slot + m, i, j, k) = F(b, fd->slot + m, ii, jj, kk);
            }
}

static void fill_const(OBlock *b, const FDesc *fd, int dir, const double *val) {
    int ilo[3], ihi[3], lo[3], hi[3];
    inner_slab(b, fd->loc, dir, ilo, ihi);
    ghost_slab(ilo, ihi, dir, b->g, lo, hi);
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k)
                for (int m = 0; m < fd->ncomp; ++m) F(b, fd->slot + m, i, j, k) = val[m];
}

/* MercuryBoundary::BC_Farfield_Eedge_set_zerocurl (Boundary_Condition.cpp:573-623) */
static void farfield_eedge(OBlock *b, const OrcPhys *P, const FDesc *fd, int dir) {
    const double *u = P->q_pv_inf, *Bi = P->B_imf;
    double Ei[3];
    Ei[0] = Bi[1] * u[2] - Bi[2] * u[1];
    Ei[1] = Bi[2] * u[0] - Bi[0] * u[2];
    Ei[2] = Bi[0] * u[1] - Bi[1] * u[0];
    int a = fd->loc - L_EX;
    int ilo[3], ihi[3];
    inner_slab(b, fd->loc, dir, ilo, ihi);
    for (int i = ilo[0]; i < ihi[0]; ++i)
        for (int j = ilo[1]; j < ihi[1]; ++j)
            for (int k = ilo[2]; k < ihi[2]; ++k) {
                int ir = i + (a == 0), jr = j + (a == 1), kr = k + (a == 2);
                F(b, fd->slot, i, j, k) = Ei[0] * (XN(b, ir, jr, kr) - XN(b, i, j, k)) + Ei[1] * (YN(b, ir, jr, kr) - YN(b, i, j, k)) +
                                          Ei[2] * (ZN(b, ir, jr, kr) - ZN(b, i, j, k));
            }
    default_copy(b, fd, dir);
}

/* the far-field Eres handler with a zero far field: boundary-layer edges <- 0 */
static void farfield_eedge_zero(OBlock *b, const FDesc *fd, int dir) {
    int ilo[3], ihi[3];
    inner_slab(b, fd->loc, dir, ilo, ihi);
    for (int i = ilo[0]; i < ihi[0]; ++i)
        for (int j = ilo[1]; j < ihi[1]; ++j)
            for (int k = ilo[2]; k < ihi[2]; ++k) F(b, fd->slot, i, j, k) = 0.0;
    default_copy(b, fd, dir);
}

/* MercuryBoundary::BC_Solid_Surface_ (Boundary_Condition_Fluid.cpp:36-260):
 * non-Na species: first active layer <- (rho, p) one cell deeper, zero velocity;
 * ghosts <- first layer.  Na species: ghosts <- floor state. */
static void wall_fluid(OBlock *b, const OrcPhys *P, int slot, int species, int dir) {
    int ax = abs(dir), sgn = dir > 0 ? 1 : -1;
    int cyc[3] = {ax == 1 ? sgn : 0, ax == 2 ? sgn : 0, ax == 3 ? sgn : 0};
    /* He++ takes the H+ rule with its own density floor (no reference counterpart) */
    const int is_na = species == 1;
    const double rho_floor = species == 1 ? P->wall_rho_floor_Na : (species == 2 ? P->wall_rho_floor_He : P->wall_rho_floor_H);
    const double p_floor = P->wall_p_floor;
    const double gm1 = P->gamma - 1.0;
    int ilo[3], ihi[3];
    inner_slab(b, L_CELL, dir, ilo, ihi);
    int ulo[3], uhi[3];
    alloc_box(b, L_CELL, b->g, ulo, uhi);
#define CLAMP(v, a) ((v) < ulo[a] ? ulo[a] : ((v) > uhi[a] - 1 ? uhi[a] - 1 : (v)))
    for (int ii = ilo[0]; ii < ihi[0]; ++ii)
        for (int jj = ilo[1]; jj < ihi[1]; ++jj)
            for (int kk = ilo[2]; kk < ihi[2]; ++kk) {
                if (is_na) {
                    for (int ng = 1; ng <= b->g; ++ng) {
                        int ig = CLAMP(ii + ng * cyc[0], 0), jg = CLAMP(jj + ng * cyc[1], 1), kg = CLAMP(kk + ng * cyc[2], 2);
                        F(b, slot + 0, ig, jg, kg) = rho_floor;
                        F(b, slot + 1, ig, jg, kg) = 0.0;
                        F(b, slot + 2, ig, jg, kg) = 0.0;
                        F(b, slot + 3, ig, jg, kg) = 0.0;
                        F(b, slot + 4, ig, jg, kg) = p_floor / gm1;
                    }
                    continue;
                }
                for (int layer = 0; layer <= b->g; ++layer) {
                    int ir, jr, kr, ig, jg, kg;
                    if (layer == 0) {
                        ir = CLAMP(ii - cyc[0], 0); jr = CLAMP(jj - cyc[1], 1); kr = CLAMP(kk - cyc[2], 2);
                        ig = ii; jg = jj; kg = kk;
                    } else {
                        ir = CLAMP(ii, 0); jr = CLAMP(jj, 1); kr = CLAMP(kk, 2);
                        ig = CLAMP(ii + layer * cyc[0], 0); jg = CLAMP(jj + layer * cyc[1], 1); kg = CLAMP(kk + layer * cyc[2], 2);
                    }
                    /* pressure_from_cons (Boundary_Condition_Fluid.cpp:79-93) */
                    double rho_r = dmax(F(b, slot, ir, jr, kr), rho_floor);
                    double mxr = F(b, slot + 1, ir, jr, kr), myr = F(b, slot + 2, ir, jr, kr), mzr = F(b, slot + 3, ir, jr, kr);
                    double Er = F(b, slot + 4, ir, jr, kr);
                    double rho2 = dmax(rho_r, rho_floor);
                    double ke = 0.5 * (mxr * mxr + myr * myr + mzr * mzr) / rho2;
                    double eint = Er - ke;
                    if (eint < 0.0) eint = 0.0;
                    double p_r = dmax(gm1 * eint, p_floor);
                    F(b, slot + 0, ig, jg, kg) = rho_r;
                    F(b, slot + 1, ig, jg, kg) = rho_r * 0.0;
                    F(b, slot + 2, ig, jg, kg) = rho_r * 0.0;
                    F(b, slot + 3, ig, jg, kg) = rho_r * 0.0;
                    const double ke_g = 0.5 * rho_r * 0.0;
                    F(b, slot + 4, ig, jg, kg) = p_r / gm1 + ke_g;
                }
            }
#undef CLAMP
}

/* MercuryBoundary::BC_Solid_Surface_Jcell (Boundary_Condition_EM.cpp:382-427) */
static void wall_jcell(OBlock *b, const FDesc *fd, int dir) {
    int ax = abs(dir), sgn = dir > 0 ? 1 : -1, inward = -sgn;
    int ilo[3], ihi[3];
    inner_slab(b, L_CELL, dir, ilo, ihi);
    for (int i = ilo[0]; i < ihi[0]; ++i)
        for (int j = ilo[1]; j < ihi[1]; ++j)
            for (int k = ilo[2]; k < ihi[2]; ++k) {
                for (int n = 0; n < 3; ++n) F(b, fd->slot + n, i, j, k) = 0.0;
                int is = i + (ax == 1 ? inward : 0), js = j + (ax == 2 ? inward : 0), ks = k + (ax == 3 ? inward : 0);
                for (int n = 0; n < 3; ++n) F(b, fd->slot + n, is, js, ks) = 0.0;
            }
    default_copy(b, fd, dir);
}

enum { G_UCELL, G_BFACE, G_EEDGE, G_EHALL, G_JEDGE, G_BCELL, G_JCELL, G_BADD, G_ERES, G_ERES0, G_DB, G_DJ, G_NGROUP };
/* field kinds for handler dispatch (K_ERES0: Eres with a zero far-field
 * field, the linear part of the resistive operator) */
enum { K_UH, K_UNA, K_BFACE, K_EEDGE, K_EHALL, K_JEDGE, K_BCELL, K_BIND, K_JCELL, K_BADD, K_ERES, K_ERES0, K_DB, K_DJ, K_UHE };

typedef struct {
    int kind, slot, ncomp, loc, level; /* level: 1 = Corner1D, 3 = Corner3D */
    int owner_sync;
} GField;

static int group_fields(const OrcPhys *P, int gid, GField out[6]) {
    int n = 0;
#define ADD(K, S, NC, L, LV, OS) out[n].kind = K, out[n].slot = S, out[n].ncomp = NC, out[n].loc = L, out[n].level = LV, out[n].owner_sync = OS, ++n
    switch (gid) {
    case G_UCELL:
        ADD(K_UH, S_UH, 5, L_CELL, 1, 0);
        ADD(K_UNA, S_UN, 5, L_CELL, 1, 0);
        if (n_species(P) > 2) ADD(K_UHE, S_UHE, 5, L_CELL, 1, 0);
        break;
    case G_BFACE: for (int a = 0; a < 3; ++a) ADD(K_BFACE, S_B + a, 1, L_FX + a, 3, 1); break;
    case G_BADD: for (int a = 0; a < 3; ++a) ADD(K_BADD, S_BADD + a, 1, L_FX + a, 3, 1); break;
    case G_EEDGE: for (int a = 0; a < 3; ++a) ADD(K_EEDGE, S_E + a, 1, L_EX + a, 3, 1); break;
    case G_EHALL: for (int a = 0; a < 3; ++a) ADD(K_EHALL, S_EH + a, 1, L_EX + a, 3, 1); break;
    case G_JEDGE: for (int a = 0; a < 3; ++a) ADD(K_JEDGE, S_J + a, 1, L_EX + a, 3, 1); break;
    case G_BCELL: ADD(K_BCELL, S_BC, 3, L_CELL, 3, 0); ADD(K_BIND, S_BI, 3, L_CELL, 3, 0); break;
    case G_JCELL: ADD(K_JCELL, S_JC, 3, L_CELL, 3, 0); break;
    /* Eres / dB / dJ groups (Boundary_Build.cpp:605-613, 674-693) */
    case G_ERES: for (int a = 0; a < 3; ++a) ADD(K_ERES, S_ERES + a, 1, L_EX + a, 3, 1); break;
    case G_ERES0: for (int a = 0; a < 3; ++a) ADD(K_ERES0, S_ERES + a, 1, L_EX + a, 3, 1); break;
    case G_DB: for (int a = 0; a < 3; ++a) ADD(K_DB, S_DB + a, 1, L_FX + a, 3, 1); break;
    case G_DJ: for (int a = 0; a < 3; ++a) ADD(K_DJ, S_DJ + a, 1, L_EX + a, 3, 1); break;
    }
#undef ADD
    return n;
}

/* one physical region handler (Z4_Mercury/src/boundary/Boundary_Build.cpp:5-300) */
static void apply_handler(OBlock *b, const OrcPhys *P, const GField *gf, int bc, int dir) {
    FDesc fd = {gf->slot, gf->ncomp, gf->loc};
    static const double zero3[3] = {0, 0, 0};
    if (bc == BC_FARFIELD) {
        switch (gf->kind) {
        case K_UH: fill_const(b, &fd, dir, P->qinf); break;
        case K_UNA: fill_const(b, &fd, dir, P->qinfs); break;
        case K_UHE: fill_const(b, &fd, dir, P->qinf_He); break;
        case K_BFACE: break; /* nop */
        case K_BCELL: fill_const(b, &fd, dir, P->B_imf); break;
        case K_BIND: fill_const(b, &fd, dir, zero3); break;
        case K_EEDGE: farfield_eedge(b, P, &fd, dir); break;
        case K_ERES: farfield_eedge(b, P, &fd, dir); break; /* Boundary_Build.cpp:198-200 */
        case K_ERES0: farfield_eedge_zero(b, &fd, dir); break;
        default: default_copy(b, &fd, dir); break; /* Ehall, J edges, J_cell, Badd, dB, dJ */
        }
    } else if (bc == BC_WALL) {
        switch (gf->kind) {
        case K_UH: wall_fluid(b, P, gf->slot, 0, dir); break;
        case K_UNA: wall_fluid(b, P, gf->slot, 1, dir); break;
        case K_UHE: wall_fluid(b, P, gf->slot, 2, dir); break;
        case K_JCELL: wall_jcell(b, &fd, dir); break;
        default: break; /* nop for every other field on the fluid side of the wall */
        }
    }
}

/* connection of block face f: 1 and (*nb, off) for a periodic self-connection
 * (translation by the block size) or an interface, 0 for a physical face */
static int face_conn(const Orc *o, int ib, int f, int *nb, int off[3]) {
    const OBlock *b = &o->b[ib];
    const int a = f / 2;
    off[0] = off[1] = off[2] = 0;
    if (b->bc[f] == BC_PERIODIC) {
        *nb = ib;
        off[a] = (f & 1) ? -b->m[a] : b->m[a];
        return 1;
    }
    if (b->bc[f] == BC_INTERFACE) {
        *nb = b->nbr[f];
        for (int c = 0; c < 3; ++c) off[c] = b->off[f][c];
        return 1;
    }
    return 0;
}

/* fill_physical_corner_ghost (src/5_boundary/Boundary_Physical.cpp:39-59) on the
 * corner ghost boxes of physical edges (2 physical faces) and vertices (3):
 * an edge or vertex with a periodic or interface face is an inner patch
 * (src/2_topology/TopologyEdge.cpp:161-166) and gets no corner default. */
static void corner_default(OBlock *b, const GField *gf) {
    int ohi[3];
    owned_hi(b, gf->loc, ohi);
    int lo[3], hi[3];
    alloc_box(b, gf->loc, b->g, lo, hi);
    /* writes corner ghosts only, reads owned points only */
    PAR2
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                int idx[3] = {i, j, k}, nout = 0, ok = 1;
                for (int a = 0; a < 3; ++a) {
                    int side = idx[a] < 0 ? 0 : (idx[a] >= ohi[a] ? 1 : -1);
                    if (side < 0) continue;
                    ++nout;
                    const int c = b->bc[2 * a + side];
                    if (c != BC_FARFIELD && c != BC_WALL) ok = 0;
                }
                if (nout < 2 || !ok) continue;
                int ii = i < 0 ? 0 : (i > ohi[0] - 1 ? ohi[0] - 1 : i);
                int jj = j < 0 ? 0 : (j > ohi[1] - 1 ? ohi[1] - 1 : j);
                int kk = k < 0 ? 0 : (k > ohi[2] - 1 ? ohi[2] - 1 : k);
                for (int m = 0; m < gf->ncomp; ++m) F(b, gf->slot + m, i, j, k) = F(b, gf->slot + m, ii, jj, kk);
            }
}

/* One stage of the staged halo (MercuryBoundary::Sync_,
 * Z4_Mercury/src/boundary/Boundary_Sync_.cpp:11-99; Halo::sync_group,
 * src/4_halo/core/HaloSyncFramework.cpp:969-1035) for the ghost DOFs with
 * exactly `stage` out-of-range axes: stage 1 the face strips
 * (build_inner_1DCorner_pattern, src/4_halo/build/HaloBuildFace.cpp:6-57; strips
 * of LayoutTraits.cpp:93-130, the shared boundary DOF excluded on the source
 * side), stage 2 the edge and stage 3 the vertex corner boxes.  A ghost is
 * filled when one of its out-of-range faces is a periodic or interface
 * connection (the inner patch wins over physical ones,
 * src/2_topology/TopologyEdge.cpp:161-166): its value is read one hop away,
 * at the translated index in the neighbour across the first such face (axis
 * order), copied as in exchange_inner / exchange_inner_edge / _vertex
 * (src/4_halo/component_copy/HaloExchangeInner.cpp:90-247).  The hop removes
 * one out-of-range axis, so a stage reads owned values or ghosts that an earlier
 * stage, a physical handler or the corner default has written, never its own
 * destinations. */
static void halo_stage(Orc *o, int ib, const GField *gf, int stage) {
    OBlock *b = &o->b[ib];
    int ohi[3], lo[3], hi[3];
    owned_hi(b, gf->loc, ohi);
    alloc_box(b, gf->loc, b->g, lo, hi);
    PAR2
    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k) {
                int idx[3] = {i, j, k}, nout = 0, hop = -1;
                for (int a = 0; a < 3; ++a) {
                    if (idx[a] >= 0 && idx[a] < ohi[a]) continue;
                    ++nout;
                    const int c = b->bc[2 * a + (idx[a] < 0 ? 0 : 1)];
                    if (hop < 0 && (c == BC_PERIODIC || c == BC_INTERFACE)) hop = a;
                }
                if (nout != stage || hop < 0) continue;
                int nb, off[3];
                face_conn(o, ib, 2 * hop + (idx[hop] < 0 ? 0 : 1), &nb, off);
                OBlock *s = &o->b[nb];
                for (int m = 0; m < gf->ncomp; ++m)
                    F(b, gf->slot + m, i, j, k) = F(s, gf->slot + m, i + off[0], j + off[1], k + off[2]);
            }
}

/* owner -> alias (Halo::sync_owner_alias_group, HaloSyncFramework.cpp:695-712,
 * pattern built at :501-560): every alias of a class of shared DOFs whose axis
 * is the field's takes sign(owner) * sign(alias) times the owner's value, read
 * from the triplet field of the owner's axis.  Classes of one dimension may sit
 * under any of its locations, so all three are scanned. */
static void owner_sync(Orc *o, const GField *gf) {
    const int Lb = gf->loc >= L_EX && gf->loc <= L_EZ ? L_EX : (gf->loc >= L_FX && gf->loc <= L_FZ ? L_FX : gf->loc);
    const int a = gf->loc - Lb, base = gf->slot - a, nL = Lb == L_NODE || Lb == L_CELL ? 1 : 3;
    for (int L = Lb; L < Lb + nL; ++L)
        for (int c = 0; c < o->ncls[L]; ++c) {
            const long s0 = o->cls_start[L][c], s1 = o->cls_start[L][c + 1];
            const double v = o->b[o->cls_b[L][s0]].f[base + o->cls_ax[L][s0]][o->cls_q[L][s0]];
            for (long t = s0 + 1; t < s1; ++t) {
                if (o->cls_ax[L][t] != a) continue;
                const int sg = o->cls_sg[L][s0] * o->cls_sg[L][t];
                o->b[o->cls_b[L][t]].f[gf->slot][o->cls_q[L][t]] = sg * v;
            }
        }
}

/* ------------------------------------------------------------------------- */
/* patch connections: rotated, mirrored and partial-face block interfaces     */
/* ------------------------------------------------------------------------- */
/* A patch joins a node box on one block face to a node box of a neighbour face
 * by an IndexTransform (include/2_topology/TopologyTypes.h:93-98): neighbour
 * node [perm[a]] = sign[a] * node[a] + offset[a].  Record of 18 ints. */
enum { PT_BLOCK = 0, PT_DIR = 1, PT_LO = 2, PT_HI = 5, PT_NB = 8, PT_PERM = 9, PT_SIGN = 12, PT_OFF = 15, PT_N = 18 };
/* A halo region of one level: receiving and sending block, node box on the
 * receiver, the out-of-range directions of its ghost box, the transform, and
 * whether the source is found by the inverse map (face level: the transform
 * runs sender -> receiver) or the forward map (edge and vertex levels:
 * receiver -> sender).  Record of 22 ints. */
enum { RG_RECV = 0, RG_SEND = 1, RG_LO = 2, RG_HI = 5, RG_ND = 8, RG_DIR = 9, RG_PERM = 12, RG_SIGN = 15, RG_OFF = 18, RG_INV = 21, RG_N = 22 };

/* node index map of a transform (HALO_TOOLS::apply_transform, src/4_halo/build/HaloBuildTools.cpp:296-310) */
static void tr_map(const int *perm, const int *sign, const int *off, const int p[3], int out[3]) {
    for (int a = 0; a < 3; ++a) out[perm[a]] = sign[a] * p[a] + off[a];
}
/* TOPO::transform_node_box (src/2_topology/TopologyOps.cpp:91-117) */
static void tr_box(const int *perm, const int *sign, const int *off, const int lo[3], const int hi[3], int olo[3], int ohi[3]) {
    const int h[3] = {hi[0] - 1, hi[1] - 1, hi[2] - 1};
    int p0[3], p1[3];
    tr_map(perm, sign, off, lo, p0);
    tr_map(perm, sign, off, h, p1);
    for (int a = 0; a < 3; ++a) {
        olo[a] = p0[a] < p1[a] ? p0[a] : p1[a];
        ohi[a] = (p0[a] < p1[a] ? p1[a] : p0[a]) + 1;
    }
}
/* block faces a node box lies on (boundary_dir_count, src/4_halo/build/HaloBuildEdge.cpp:9-25) */
static int bdir_count(const int lo[3], const int hi[3], const int m[3]) {
    int c = 0;
    for (int a = 0; a < 3; ++a) c += (lo[a] == 0 && hi[a] == 1) + (lo[a] == m[a] && hi[a] == m[a] + 1);
    return c;
}
/* HALO_TOOLS::detect_face_direction (HaloBuildTools.cpp:12-37); 0 if the box is on no face */
static int face_dir_of(const int lo[3], const int hi[3], const int m[3]) {
    for (int a = 0; a < 3; ++a) {
        if (lo[a] == 0 && hi[a] == 1) return -(a + 1);
        if (lo[a] == m[a] && hi[a] == m[a] + 1) return a + 1;
    }
    return 0;
}
static int inv_perm_axis(const int *perm, int d) {
    for (int a = 0; a < 3; ++a)
        if (perm[a] == d) return a;
    return d;
}
static int perm_parity(const int *perm) {
    int inv = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) inv += perm[a] > perm[b];
    return inv & 1 ? -1 : 1;
}

static void push_region(Orc *o, int lvl, int recv, int send, const int lo[3], const int hi[3], int nd, const int *dirs,
                        const int *perm, const int *sign, const int *off, int inv) {
    o->reg[lvl] = (int *)realloc(o->reg[lvl], sizeof(int) * RG_N * (o->nreg[lvl] + 1));
    int *r = o->reg[lvl] + RG_N * o->nreg[lvl]++;
    r[RG_RECV] = recv;
    r[RG_SEND] = send;
    for (int a = 0; a < 3; ++a) {
        r[RG_LO + a] = lo[a];
        r[RG_HI + a] = hi[a];
        r[RG_DIR + a] = a < nd ? dirs[a] : 0;
        r[RG_PERM + a] = perm[a];
        r[RG_SIGN + a] = sign[a];
        r[RG_OFF + a] = off[a];
    }
    r[RG_ND] = nd;
    r[RG_INV] = inv;
}

/* a face patch or a physical face of one block, and an edge patch (two faces
 * of one block), as the reference's topology lists them */
typedef struct { int inner, dir, lo[3], hi[3], nb, perm[3], sign[3], off[3]; } OFacePatch;
typedef struct { int inner, block, lo[3], hi[3], dir1, dir2, nb, perm[3], sign[3], off[3]; } OEdgePatch;

/* Halo regions of the three levels:
 * level 1 from every patch (Halo::build_inner_1DCorner_pattern,
 *   src/4_halo/build/HaloBuildFace.cpp:6-57): the receiver is the neighbour,
 *   its ghost strip outside the neighbour's face box;
 * level 2 from the edge patches (TOPO::build_edge_patches,
 *   src/2_topology/TopologyEdge.cpp:45-211): every pair of faces of one block
 *   (patches in list order, then physical faces) on different axes whose node
 *   boxes meet; the inner face owns the edge over a physical one, else the
 *   first of the pair; an inner edge becomes a region when it lies on two faces
 *   of both blocks (build_inner_2DCorner_pattern, HaloBuildEdge.cpp:188-265);
 * level 3 from the vertex patches (TOPO::build_vertex_patches,
 *   TopologyEdge.cpp:213-405; build_inner_3DCorner_pattern,
 *   src/4_halo/build/HaloBuildVertex.cpp:190-275): pairs of edges of one block
 *   on different axes meeting in one node, the first pair per node, owned by
 *   the first inner edge among all edges touching the node. */
static void build_patch_regions(Orc *o) {
    for (int l = 0; l < 4; ++l) {
        free(o->reg[l]);
        o->reg[l] = NULL;
        o->nreg[l] = 0;
    }
    for (int p = 0; p < o->npatch; ++p) {
        const int *P = o->patch + PT_N * p;
        const OBlock *nb = &o->b[P[PT_NB]];
        int lo[3], hi[3];
        tr_box(P + PT_PERM, P + PT_SIGN, P + PT_OFF, P + PT_LO, P + PT_HI, lo, hi);
        const int d = face_dir_of(lo, hi, nb->m);
        push_region(o, 1, P[PT_NB], P[PT_BLOCK], lo, hi, 1, &d, P + PT_PERM, P + PT_SIGN, P + PT_OFF, 1);
    }
    OEdgePatch *ie = NULL, *pe = NULL;
    int nie = 0, npe = 0;
    for (int ib = 0; ib < o->nb; ++ib) {
        const OBlock *b = &o->b[ib];
        OFacePatch *fl = (OFacePatch *)calloc(o->npatch + 6, sizeof(OFacePatch));
        int nf = 0;
        for (int p = 0; p < o->npatch; ++p) {
            const int *P = o->patch + PT_N * p;
            if (P[PT_BLOCK] != ib) continue;
            OFacePatch *f = &fl[nf++];
            f->inner = 1;
            f->dir = P[PT_DIR];
            f->nb = P[PT_NB];
            for (int a = 0; a < 3; ++a) {
                f->lo[a] = P[PT_LO + a];
                f->hi[a] = P[PT_HI + a];
                f->perm[a] = P[PT_PERM + a];
                f->sign[a] = P[PT_SIGN + a];
                f->off[a] = P[PT_OFF + a];
            }
        }
        for (int fc = 0; fc < 6; ++fc) {
            if (b->bc[fc] != BC_FARFIELD && b->bc[fc] != BC_WALL) continue;
            OFacePatch *f = &fl[nf++];
            const int a = fc / 2;
            f->inner = 0;
            f->dir = (fc & 1) ? a + 1 : -(a + 1);
            f->nb = -1;
            for (int c = 0; c < 3; ++c) {
                f->lo[c] = 0;
                f->hi[c] = b->m[c] + 1;
            }
            f->lo[a] = (fc & 1) ? b->m[a] : 0;
            f->hi[a] = f->lo[a] + 1;
        }
        for (int i = 0; i < nf; ++i)
            for (int j = i + 1; j < nf; ++j) {
                const OFacePatch *f1 = &fl[i], *f2 = &fl[j];
                if (abs(f1->dir) == abs(f2->dir)) continue;
                OEdgePatch e;
                int empty = 0;
                for (int a = 0; a < 3; ++a) {
                    e.lo[a] = f1->lo[a] > f2->lo[a] ? f1->lo[a] : f2->lo[a];
                    e.hi[a] = f1->hi[a] < f2->hi[a] ? f1->hi[a] : f2->hi[a];
                    empty |= e.hi[a] <= e.lo[a];
                }
                if (empty) continue;
                const OFacePatch *own = f2->inner > f1->inner ? f2 : f1;
                e.inner = own->inner;
                e.block = ib;
                e.nb = own->nb;
                for (int a = 0; a < 3; ++a) {
                    e.perm[a] = own->perm[a];
                    e.sign[a] = own->sign[a];
                    e.off[a] = own->off[a];
                }
                e.dir1 = own == f1 ? f1->dir : f2->dir;
                e.dir2 = own == f1 ? f2->dir : f1->dir;
                if (e.inner) {
                    ie = (OEdgePatch *)realloc(ie, sizeof(OEdgePatch) * (nie + 1));
                    ie[nie++] = e;
                } else {
                    pe = (OEdgePatch *)realloc(pe, sizeof(OEdgePatch) * (npe + 1));
                    pe[npe++] = e;
                }
            }
        free(fl);
    }
    for (int t = 0; t < nie; ++t) {
        const OEdgePatch *e = &ie[t];
        int lo[3], hi[3];
        tr_box(e->perm, e->sign, e->off, e->lo, e->hi, lo, hi);
        if (bdir_count(e->lo, e->hi, o->b[e->block].m) != 2 || bdir_count(lo, hi, o->b[e->nb].m) != 2) continue;
        const int dirs[2] = {e->dir1, e->dir2};
        push_region(o, 2, e->block, e->nb, e->lo, e->hi, 2, dirs, e->perm, e->sign, e->off, 0);
    }
    for (int ib = 0; ib < o->nb; ++ib) {
        const OEdgePatch **el = (const OEdgePatch **)calloc(nie + npe + 1, sizeof(OEdgePatch *));
        int ne = 0;
        for (int t = 0; t < nie; ++t)
            if (ie[t].block == ib) el[ne++] = &ie[t];
        for (int t = 0; t < npe; ++t)
            if (pe[t].block == ib) el[ne++] = &pe[t];
        int *used = (int *)malloc(sizeof(int) * 3 * (ne * ne + 1));
        int nused = 0;
        for (int i = 0; i < ne; ++i)
            for (int j = i + 1; j < ne; ++j) {
                const OEdgePatch *e1 = el[i], *e2 = el[j];
                if (6 - abs(e1->dir1) - abs(e1->dir2) == 6 - abs(e2->dir1) - abs(e2->dir2)) continue;
                int lo[3], hi[3], unit = 1;
                for (int a = 0; a < 3; ++a) {
                    lo[a] = e1->lo[a] > e2->lo[a] ? e1->lo[a] : e2->lo[a];
                    hi[a] = e1->hi[a] < e2->hi[a] ? e1->hi[a] : e2->hi[a];
                    unit &= hi[a] - lo[a] == 1;
                }
                if (!unit) continue;
                int dup = 0;
                for (int u = 0; u < nused && !dup; ++u) dup = used[3 * u] == lo[0] && used[3 * u + 1] == lo[1] && used[3 * u + 2] == lo[2];
                if (dup) continue;
                for (int a = 0; a < 3; ++a) used[3 * nused + a] = lo[a];
                ++nused;
                /* candidates: e1, e2, then every other edge touching the node */
                const OEdgePatch **cand = (const OEdgePatch **)calloc(ne + 1, sizeof(OEdgePatch *));
                int nc = 0;
                cand[nc++] = e1;
                cand[nc++] = e2;
                for (int k = 0; k < ne; ++k) {
                    if (k == i || k == j) continue;
                    int meet = 1;
                    for (int a = 0; a < 3; ++a) meet &= el[k]->lo[a] <= lo[a] && lo[a] < el[k]->hi[a];
                    if (meet) cand[nc++] = el[k];
                }
                const OEdgePatch *own = cand[0];
                for (int k = 1; k < nc; ++k)
                    if (cand[k]->inner > own->inner) own = cand[k];
                int dba[4] = {0, 0, 0, 0};
                for (int k = 0; k < nc; ++k) {
                    const int d[2] = {cand[k]->dir1, cand[k]->dir2};
                    for (int q = 0; q < 2; ++q)
                        if (!dba[abs(d[q])]) dba[abs(d[q])] = d[q];
                }
                free(cand);
                if (!own->inner) continue;
                int nlo[3], nhi[3];
                tr_box(own->perm, own->sign, own->off, lo, hi, nlo, nhi);
                if (bdir_count(lo, hi, o->b[ib].m) != 3 || bdir_count(nlo, nhi, o->b[own->nb].m) != 3) continue;
                push_region(o, 3, ib, own->nb, lo, hi, 3, dba + 1, own->perm, own->sign, own->off, 0);
            }
        free(used);
        free(el);
    }
    free(ie);
    free(pe);
}

/* ghost box of a region at location L: the DOF box of its node box, moved
 * outside along each of its directions (LAYOUT::node_box_to_dof_box and
 * apply_ghost_strip, src/0_basic/LayoutTraits.cpp:104-118, 296-303) */
static void region_box(const int *r, int L, int G, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = r[RG_LO + a];
        hi[a] = r[RG_HI + a] - DELTA[L][a];
    }
    for (int t = 0; t < r[RG_ND]; ++t) {
        const int d = r[RG_DIR + t], ax = abs(d) - 1;
        if (d < 0) {
            lo[ax] = -G;
            hi[ax] = 0;
        } else {
            lo[ax] = hi[ax];
            hi[ax] = lo[ax] + G;
        }
    }
}
/* source index of receiver point p (location Lr) of a region.  Rule 0 maps the
 * DOF's indices as node indices, as the reference does for every location
 * (HaloExchangeInner.cpp:35-52, HaloTriplet.cpp:53-84); rule 1 maps its
 * position: along a reversed axis where the DOF sits between nodes, the
 * mirrored DOF is one index lower. */
static void region_src(int rule, const int *r, int Lr, const int p[3], int s[3]) {
    const int *perm = r + RG_PERM, *sign = r + RG_SIGN, *off = r + RG_OFF;
    if (r[RG_INV]) {
        for (int a = 0; a < 3; ++a) {
            s[a] = sign[a] * (p[perm[a]] - off[a]);
            if (rule && sign[a] < 0) s[a] -= DELTA[Lr][perm[a]];
        }
    } else {
        for (int a = 0; a < 3; ++a) {
            s[perm[a]] = sign[a] * p[a] + off[a];
            if (rule && sign[a] < 0) s[perm[a]] -= DELTA[Lr][a];
        }
    }
}
static int in_alloc(const OBlock *b, int L, const int s[3]) {
    int lo[3], hi[3];
    alloc_box(b, L, b->g, lo, hi);
    for (int a = 0; a < 3; ++a)
        if (s[a] < lo[a] || s[a] >= hi[a]) return 0;
    return 1;
}

/* component copy of one level (Halo::exchange_inner / _edge / _vertex,
 * src/4_halo/component_copy/HaloExchangeInner.cpp:184-247): every region's
 * values are read first, then written in region order */
static void patch_exchange_component(Orc *o, const GField *gf, int lvl) {
    long n = 0;
    for (int t = 0; t < o->nreg[lvl]; ++t) {
        int lo[3], hi[3];
        region_box(o->reg[lvl] + RG_N * t, gf->loc, o->b[0].g, lo, hi);
        n += (long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    }
    int *db = (int *)malloc(sizeof(int) * (n + 1));
    long *dq = (long *)malloc(sizeof(long) * (n + 1));
    double *v = (double *)malloc(sizeof(double) * gf->ncomp * (n + 1));
    long u = 0;
    for (int t = 0; t < o->nreg[lvl]; ++t) {
        const int *r = o->reg[lvl] + RG_N * t;
        OBlock *rb = &o->b[r[RG_RECV]], *sb = &o->b[r[RG_SEND]];
        int lo[3], hi[3], p[3], s[3];
        region_box(r, gf->loc, rb->g, lo, hi);
        for (p[0] = lo[0]; p[0] < hi[0]; ++p[0])
            for (p[1] = lo[1]; p[1] < hi[1]; ++p[1])
                for (p[2] = lo[2]; p[2] < hi[2]; ++p[2]) {
                    region_src(o->rule, r, gf->loc, p, s);
                    if (!in_alloc(sb, gf->loc, s)) continue;
                    db[u] = r[RG_RECV];
                    dq[u] = IX(rb, p[0], p[1], p[2]);
                    for (int m = 0; m < gf->ncomp; ++m) v[gf->ncomp * u + m] = F(sb, gf->slot + m, s[0], s[1], s[2]);
                    ++u;
                }
    }
    for (long t = 0; t < u; ++t) {
        for (int m = 0; m < gf->ncomp; ++m) o->b[db[t]].f[gf->slot + m][dq[t]] = v[gf->ncomp * t + m];
        if (o->wm[0][db[t]]) o->wm[0][db[t]][dq[t]] = 1;
    }
    free(db);
    free(dq);
    free(v);
}

/* face 2-form or edge 1-form triplet of one level, written as it is read
 * (Halo::exchange_inner_{face,edge,vertex}_{face_2form,edge_1form}_triplet_,
 * src/4_halo/forms/HaloTriplet.cpp:355-489, 641-829): the receiver's axis-d
 * field takes factor * the sender's field of the corresponding axis.
 * Face level (transform sender -> receiver): source axis a with perm[a] = d;
 * factor sign[a] for 1-forms and sign[0]*sign[1]*sign[2]*sign[a] for 2-forms
 * (face_2form_orientation_sign, HaloBuildTools.cpp:312-327).  Edge and vertex
 * levels (receiver -> sender): 1-forms read axis perm[d] with factor sign[d];
 * 2-forms read axis a with perm[a] = d with the same 2-form factor.  Rule 1
 * instead carries the area vector: factor parity(perm) * sign[0]*sign[1]*
 * sign[2] * sign of the axis mapped, and 2-forms at the corner levels read axis
 * perm[d] (sign[d]). */
static void patch_exchange_triplet(Orc *o, int slot0, int face, int lvl) {
    const int Lb = face ? L_FX : L_EX;
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < o->nreg[lvl]; ++t) {
            const int *r = o->reg[lvl] + RG_N * t, *perm = r + RG_PERM, *sign = r + RG_SIGN;
            const int det = sign[0] * sign[1] * sign[2], par = o->rule ? perm_parity(perm) : 1;
            int sa, fac;
            if (lvl == 1) {
                sa = inv_perm_axis(perm, d);
                fac = face ? par * det * sign[sa] : sign[sa];
            } else if (!face) {
                sa = perm[d];
                fac = sign[d];
            } else if (o->rule) {
                sa = perm[d];
                fac = par * det * sign[d];
            } else {
                sa = inv_perm_axis(perm, d);
                fac = det * sign[sa];
            }
            OBlock *rb = &o->b[r[RG_RECV]], *sb = &o->b[r[RG_SEND]];
            int lo[3], hi[3], p[3], s[3];
            region_box(r, Lb + d, rb->g, lo, hi);
            for (p[0] = lo[0]; p[0] < hi[0]; ++p[0])
                for (p[1] = lo[1]; p[1] < hi[1]; ++p[1])
                    for (p[2] = lo[2]; p[2] < hi[2]; ++p[2]) {
                        region_src(o->rule, r, Lb + d, p, s);
                        if (!in_alloc(sb, Lb + sa, s)) continue;
                        F(rb, slot0 + d, p[0], p[1], p[2]) = fac * F(sb, slot0 + sa, s[0], s[1], s[2]);
                        if (o->wm[d][r[RG_RECV]]) o->wm[d][r[RG_RECV]][IX(rb, p[0], p[1], p[2])] = 1;
                    }
        }
}

/* Rule 1 only (no reference counterpart): a ghost DOF that no region of its
 * level wrote takes the owned DOF at its position, found by a chain of patch
 * maps: across the block face of its first out-of-range axis, through the
 * first patch of that face whose node box holds the DOF's span along the
 * in-range tangential axes and touches the face's edge along the out-of-range
 * ones, mapped by position (region_src), until the point is owned (at most 8
 * hops).  Triplet fields carry the axis and the factor of every hop.  Returns 0
 * where no chain ends in an owned DOF (a physical face on the way). */
static int chain_source(const Orc *o, int Lb, int trip, int ib, int d0, const int p0[3], int *ob, int q[3], int *od, int *ofac) {
    int b = ib, d = d0, fac = 1, p[3] = {p0[0], p0[1], p0[2]};
    for (int hop = 0; hop < 8; ++hop) {
        const int L = trip ? Lb + d : Lb;
        const OBlock *bk = &o->b[b];
        int hi[3], a = -1;
        owned_hi(bk, L, hi);
        for (int t = 0; t < 3 && a < 0; ++t)
            if (p[t] < 0 || p[t] >= hi[t]) a = t;
        if (a < 0) {
            if (hop == 0) return 0;
            *ob = b, *od = d, *ofac = fac;
            for (int t = 0; t < 3; ++t) q[t] = p[t];
            return 1;
        }
        const int dir = p[a] < 0 ? -(a + 1) : a + 1;
        const int *P = NULL;
        for (int pp = 0; pp < o->npatch && !P; ++pp) {
            const int *C = o->patch + PT_N * pp;
            if (C[PT_BLOCK] != b || C[PT_DIR] != dir) continue;
            int ok = 1;
            for (int t = 0; t < 3 && ok; ++t) {
                if (t == a) continue;
                if (p[t] >= 0 && p[t] < hi[t]) ok = C[PT_LO + t] <= p[t] && p[t] + DELTA[L][t] <= C[PT_HI + t] - 1;
                else if (p[t] < 0) ok = C[PT_LO + t] == 0;
                else ok = C[PT_HI + t] - 1 == bk->m[t];
            }
            if (ok) P = C;
        }
        if (!P) return 0;
        const int *perm = P + PT_PERM, *sign = P + PT_SIGN, *off = P + PT_OFF;
        int np[3];
        for (int t = 0; t < 3; ++t) np[perm[t]] = sign[t] * p[t] + off[t] - (sign[t] < 0 ? DELTA[L][t] : 0);
        if (trip) {
            fac *= Lb == L_FX ? perm_parity(perm) * sign[0] * sign[1] * sign[2] * sign[d] : sign[d];
            d = perm[d];
        }
        b = P[PT_NB];
        for (int t = 0; t < 3; ++t) p[t] = np[t];
    }
    return 0;
}
static void patch_complete(Orc *o, const GField *gf, int nf, int trip, int stage) {
    const int Lb = trip ? gf[0].loc : L_CELL;
    const int nax = trip ? 3 : nf;
    for (int x = 0; x < nax; ++x) {
        const GField *f = trip ? &gf[0] : &gf[x];
        const int d = trip ? x : 0, L = trip ? Lb + x : f->loc, slot = trip ? gf[0].slot + x : f->slot, nc = trip ? 1 : f->ncomp;
        for (int ib = 0; ib < o->nb; ++ib) {
            OBlock *b = &o->b[ib];
            int lo[3], hi[3], ohi[3], p[3];
            alloc_box(b, L, b->g, lo, hi);
            owned_hi(b, L, ohi);
            for (p[0] = lo[0]; p[0] < hi[0]; ++p[0])
                for (p[1] = lo[1]; p[1] < hi[1]; ++p[1])
                    for (p[2] = lo[2]; p[2] < hi[2]; ++p[2]) {
                        int nout = 0;
                        for (int t = 0; t < 3; ++t) nout += p[t] < 0 || p[t] >= ohi[t];
                        if (nout != stage || o->wm[trip ? d : 0][ib][IX(b, p[0], p[1], p[2])]) continue;
                        int sb, q[3], sd, fac;
                        if (!chain_source(o, Lb, trip, ib, d, p, &sb, q, &sd, &fac)) continue;
                        for (int m = 0; m < nc; ++m)
                            F(b, slot + m, p[0], p[1], p[2]) = fac * F(&o->b[sb], (trip ? gf[0].slot + sd : slot) + m, q[0], q[1], q[2]);
                    }
        }
    }
}

/* MercuryBoundary::Sync_ (Z4_Mercury/src/boundary/Boundary_Sync_.cpp:4-110),
 * over all blocks phase by phase: physical handlers region by region (Farfield
 * regions before the wall, the order of the sorted physical patch list), corner
 * defaults, the halo stages up to the field's level, then owner -> alias for
 * the face/edge form groups. */
static void sync_group(Orc *o, int gid) {
    GField gf[6];
    int nf = group_fields(&o->P, gid, gf);
    for (int n = 0; n < nf; ++n)
        for (int ib = 0; ib < o->nb; ++ib) {
            OBlock *b = &o->b[ib];
            for (int pass = 0; pass < 2; ++pass) {
                int want = pass == 0 ? BC_FARFIELD : BC_WALL;
                for (int f = 0; f < 6; ++f) {
                    if (b->bc[f] != want) continue;
                    int dir = (f / 2 + 1) * ((f & 1) ? 1 : -1);
                    apply_handler(b, &o->P, &gf[n], want, dir);
                }
            }
        }
    for (int n = 0; n < nf; ++n)
        for (int ib = 0; ib < o->nb; ++ib) corner_default(&o->b[ib], &gf[n]);
    if (o->npatch > 0) {
        /* Halo::sync_group(name, stage) (HaloSyncFramework.cpp:1004-1035): the
         * group's triplet first, then its component-copy fields */
        const int trip = gf[0].loc != L_CELL;
        for (int d = 0; d < 3; ++d)
            for (int ib = 0; ib < o->nb; ++ib) o->wm[d][ib] = (char *)calloc(o->b[ib].V, 1);
        /* rule 1 completes each level before the next one reads it */
        if (trip) {
            for (int stage = 1; stage <= gf[0].level; ++stage) {
                patch_exchange_triplet(o, gf[0].slot, gf[0].loc == L_FX, stage);
                if (o->rule) patch_complete(o, gf, nf, 1, stage);
            }
        } else {
            /* the fields are independent: stage by stage per field */
            for (int n = 0; n < nf; ++n) {
                for (int ib = 0; ib < o->nb; ++ib) memset(o->wm[0][ib], 0, o->b[ib].V);
                for (int stage = 1; stage <= gf[n].level; ++stage) {
                    patch_exchange_component(o, &gf[n], stage);
                    if (o->rule) patch_complete(o, &gf[n], 1, 0, stage);
                }
            }
        }
        for (int d = 0; d < 3; ++d)
            for (int ib = 0; ib < o->nb; ++ib) {
                free(o->wm[d][ib]);
                o->wm[d][ib] = NULL;
            }
    } else {
        for (int stage = 1; stage <= 3; ++stage)
            for (int n = 0; n < nf; ++n) {
                if (stage > gf[n].level) continue;
                for (int ib = 0; ib < o->nb; ++ib) halo_stage(o, ib, &gf[n], stage);
            }
    }
    for (int n = 0; n < nf; ++n)
        if (gf[n].owner_sync) owner_sync(o, &gf[n]);
}

/* ReduceEdgeAliasCandidatesToOwners_ (MercurySolver_Current.cpp:3-57): the owner
 * of each shared edge class takes its orientation sign times the mean of
 * sign * value over all members, summed in ascending EntityKey order. */
static void alias_reduce_edges(Orc *o, int slot0) {
    for (int L = L_EX; L <= L_EZ; ++L) {
        for (int c = 0; c < o->ncls[L]; ++c) {
            const long s0 = o->cls_start[L][c], s1 = o->cls_start[L][c + 1];
            double sum = 0.0;
            for (long t = s0; t < s1; ++t) sum += o->cls_sg[L][t] * o->b[o->cls_b[L][t]].f[slot0 + o->cls_ax[L][t]][o->cls_q[L][t]];
            o->b[o->cls_b[L][s0]].f[slot0 + o->cls_ax[L][s0]][o->cls_q[L][s0]] = o->cls_sg[L][s0] * sum / (double)(s1 - s0);
        }
    }
}

/* Equivalence classes of the DOFs shared through periodic and interface faces
 * (Topology equivalence, src/2_topology/TopologyFaceEquiv.cpp,
 * TopologyEdgeEquiv.cpp; owner = smallest EntityKey): every owned DOF on a
 * connected block face is united with its translate in the neighbour.  Only
 * locations with delta = 0 along the face normal lie on the face. */
static long uf_find(long *p, long x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}
static int cmp_pair(const void *a, const void *b) {
    const long *x = (const long *)a, *y = (const long *)b;
    if (x[0] != y[0]) return (x[0] > y[0]) - (x[0] < y[0]);
    return (x[1] > y[1]) - (x[1] < y[1]);
}
static void build_classes(Orc *o) {
    int any = 0;
    for (int ib = 0; ib < o->nb; ++ib)
        for (int f = 0; f < 6; ++f) any |= o->b[ib].bc[f] == BC_PERIODIC || o->b[ib].bc[f] == BC_INTERFACE;
    for (int L = 0; L < 8; ++L) {
        o->ncls[L] = 0;
        o->cls_start[L] = (long *)calloc(1, sizeof(long));
        o->cls_b[L] = NULL;
        o->cls_q[L] = NULL;
        o->cls_ax[L] = NULL;
        o->cls_sg[L] = NULL;
    }
    if (!any) return;
    long *base = (long *)calloc(o->nb + 1, sizeof(long));
    for (int ib = 0; ib < o->nb; ++ib) base[ib + 1] = base[ib] + o->b[ib].V;
    const long tot = base[o->nb];
    long *par = (long *)malloc(sizeof(long) * tot);
    char *hit = (char *)malloc(tot);
    for (int L = L_FX; L <= L_NODE; ++L) {
        for (long x = 0; x < tot; ++x) par[x] = x;
        memset(hit, 0, tot);
        for (int ib = 0; ib < o->nb; ++ib) {
            OBlock *b = &o->b[ib];
            int ohi[3];
            owned_hi(b, L, ohi);
            for (int f = 0; f < 6; ++f) {
                const int a = f / 2;
                int nb, off[3];
                if (DELTA[L][a] != 0 || !face_conn(o, ib, f, &nb, off)) continue;
                OBlock *nbk = &o->b[nb];
                int lo[3] = {0, 0, 0}, hi[3] = {ohi[0], ohi[1], ohi[2]};
                lo[a] = (f & 1) ? b->m[a] : 0;
                hi[a] = lo[a] + 1;
                for (int i = lo[0]; i < hi[0]; ++i)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int k = lo[2]; k < hi[2]; ++k) {
                            const long x = base[ib] + IX(b, i, j, k);
                            const long y = base[nb] + IX(nbk, i + off[0], j + off[1], k + off[2]);
                            const long rx = uf_find(par, x), ry = uf_find(par, y);
                            if (rx != ry) par[rx < ry ? ry : rx] = rx < ry ? rx : ry;
                            hit[x] = hit[y] = 1;
                        }
            }
        }
        /* members grouped by class root (the smallest member, since the union
         * keeps the smaller root); the global index orders (block, i, j, k) */
        long nmem = 0;
        for (long x = 0; x < tot; ++x) nmem += hit[x];
        long *key = (long *)malloc(sizeof(long) * 2 * (nmem ? nmem : 1));
        long t = 0;
        for (long x = 0; x < tot; ++x)
            if (hit[x]) {
                key[2 * t] = uf_find(par, x);
                key[2 * t + 1] = x;
                ++t;
            }
        qsort(key, nmem, 2 * sizeof(long), cmp_pair);
        long *st = (long *)malloc(sizeof(long) * (nmem + 1));
        int *cb = (int *)malloc(sizeof(int) * (nmem ? nmem : 1));
        long *cq = (long *)malloc(sizeof(long) * (nmem ? nmem : 1));
        int nc = 0;
        for (long u = 0; u < nmem; ++u) {
            if (u == 0 || key[2 * u] != key[2 * (u - 1)]) st[nc++] = u;
            const long x = key[2 * u + 1];
            int ib = 0;
            while (base[ib + 1] <= x) ++ib;
            cb[u] = ib;
            cq[u] = x - base[ib];
        }
        st[nc] = nmem;
        free(o->cls_start[L]);
        o->ncls[L] = nc;
        o->cls_start[L] = st;
        o->cls_b[L] = cb;
        o->cls_q[L] = cq;
        o->cls_ax[L] = (int *)malloc(sizeof(int) * (nmem ? nmem : 1));
        o->cls_sg[L] = (signed char *)malloc(nmem ? nmem : 1);
        const int ax = L >= L_EX && L <= L_EZ ? L - L_EX : (L >= L_FX && L <= L_FZ ? L - L_FX : 0);
        for (long u = 0; u < nmem; ++u) {
            o->cls_ax[L][u] = ax;
            o->cls_sg[L][u] = 1;
        }
        free(key);
    }
    free(hit);
    free(par);
    free(base);
}

/* ------------------------------------------------------------------------- */
/* physics                                                                    */
/* ------------------------------------------------------------------------- */

/* n: number densities, q: charge densities Z_s n_s (q = n for the singly
 * charged H+ and Na+), chi = q / n_e (the weights of U+), w = q / ne_eff */
typedef struct {
    double n[NSMAX], q[NSMAX], ne, ne_eff, chi[NSMAX], w[NSMAX];
} NumInfo;

/* Hall_Num_Limiter (MercurySolver_Tools.cpp:111-157); with He++ its charge
 * density Z_He n_He joins n_e (no reference counterpart) */
static NumInfo num_limiter(const OrcPhys *P, const double *rho, int ns) {
    NumInfo n;
    n.n[0] = dmax(rho[0], 0.0) / P->M_H * P->M_ref;
    n.n[1] = dmax(rho[1], 0.0) / P->M_Na * P->M_ref;
    n.q[0] = n.n[0];
    n.q[1] = n.n[1];
    n.ne = n.q[0] + n.q[1];
    n.n[2] = n.q[2] = 0.0;
    if (ns > 2) {
        n.n[2] = dmax(rho[2], 0.0) / P->M_He * P->M_ref;
        n.q[2] = P->Z_He * n.n[2];
        n.ne = n.ne + n.q[2];
    }
    n.ne_eff = sqrt(n.ne * n.ne + P->ne_floor * P->ne_floor);
    if (n.ne > 1.0e-300) {
        for (int s = 0; s < ns; ++s) n.chi[s] = n.q[s] / n.ne;
    } else {
        n.chi[0] = 1.0;
        for (int s = 1; s < ns; ++s) n.chi[s] = 0.0;
    }
    for (int s = 0; s < ns; ++s) n.w[s] = n.q[s] / n.ne_eff;
    return n;
}

/* the species densities of one point */
static NumInfo num_at(const OrcPhys *P, const OBlock *b, long q) {
    const int ns = n_species(P);
    double rho[NSMAX] = {0.0, 0.0, 0.0};
    for (int s = 0; s < ns; ++s) rho[s] = b->f[SP_U[s]][q];
    return num_limiter(P, rho, ns);
}

/* calc_PV (MercurySolver_Tools.cpp:159-213), full allocated range */
static void calc_pv(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int s = 0; s < n_species(&o->P); ++s) {
            int su = SP_U[s], sp = SP_PV[s];
            double coeff = species_coeff(&o->P, s);
            PAR1
            for (long q = 0; q < b->V; ++q) {
                const double rho = dmax(b->f[su][q], 1e-12);
                const double u = b->f[su + 1][q] / rho, v = b->f[su + 2][q] / rho, w = b->f[su + 3][q] / rho;
                const double E = b->f[su + 4][q];
                const double ke = 0.5 * rho * (u * u + v * v + w * w);
                double eint = E - ke;
                if (eint < 0.0) eint = 0.0;
                double p = (o->P.gamma - 1.0) * eint;
                p = dmax(p, 1e-12);
                b->f[sp][q] = u;
                b->f[sp + 1][q] = v;
                b->f[sp + 2][q] = w;
                b->f[sp + 3][q] = p;
                b->f[sp + 4][q] = p / (rho * coeff);
            }
        }
    }
}

/* calc_Uplus (MercurySolver_Tools.cpp:215-301), full allocated range */
static void calc_uplus(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR1
        for (long q = 0; q < b->V; ++q) {
            const int ns = n_species(&o->P);
            double r[NSMAX], us[NSMAX][3];
            for (int s = 0; s < ns; ++s) {
                r[s] = dmax(b->f[SP_U[s]][q], 0.0);
                for (int c = 0; c < 3; ++c) us[s][c] = (r[s] > 1e-20) ? b->f[SP_PV[s] + c][q] : 0.0;
            }
            NumInfo n = num_limiter(&o->P, r, ns);
            for (int c = 0; c < 3; ++c) {
                double up = n.chi[0] * us[0][c] + n.chi[1] * us[1][c];
                if (ns > 2) up = up + n.chi[2] * us[2][c];
                b->f[S_UP + c][q] = up;
            }
        }
    }
}

/* calc_Bcell (MercurySolver_MagneticCell.cpp:3-81), owned cells */
static void calc_bcell(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR2
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k) {
                    double pt[6], pi[6];
                    pt[0] = -F(b, S_B + 0, i, j, k) - F(b, S_BADD + 0, i, j, k);
                    pt[1] = F(b, S_B + 0, i + 1, j, k) + F(b, S_BADD + 0, i + 1, j, k);
                    pt[2] = -F(b, S_B + 1, i, j, k) - F(b, S_BADD + 1, i, j, k);
                    pt[3] = F(b, S_B + 1, i, j + 1, k) + F(b, S_BADD + 1, i, j + 1, k);
                    pt[4] = -F(b, S_B + 2, i, j, k) - F(b, S_BADD + 2, i, j, k);
                    pt[5] = F(b, S_B + 2, i, j, k + 1) + F(b, S_BADD + 2, i, j, k + 1);
                    pi[0] = -F(b, S_B + 0, i, j, k);
                    pi[1] = F(b, S_B + 0, i + 1, j, k);
                    pi[2] = -F(b, S_B + 1, i, j, k);
                    pi[3] = F(b, S_B + 1, i, j + 1, k);
                    pi[4] = -F(b, S_B + 2, i, j, k);
                    pi[5] = F(b, S_B + 2, i, j, k + 1);
                    double bt[3] = {0, 0, 0}, bi[3] = {0, 0, 0};
                    for (int n = 0; n < 6; ++n)
                        for (int c = 0; c < 3; ++c) {
                            const double w = F(b, S_WB + 6 * c + n, i, j, k);
                            bt[c] += w * pt[n];
                            bi[c] += w * pi[n];
                        }
                    for (int c = 0; c < 3; ++c) {
                        F(b, S_BC + c, i, j, k) = bt[c];
                        F(b, S_BI + c, i, j, k) = bi[c];
                    }
                }
    }
}

/* CTOperators::CurlAdjFaceToEdge (Z4_Mercury/src/operators/CTOperators.cpp:59-110)
 * then J *= alpha (MercurySolver_Current.cpp:140-201), owned edges */
/* CTOperators::CurlAdjFaceToEdge (Z4_Mercury/src/operators/CTOperators.cpp:59-110)
 * of the face 2-form in slots sf.. times alpha, into the edge slots se.. */
static void adj_curl(Orc *o, int sf, int se) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        int hi[3];
#define BB(a, i, j, k) (F(b, S_BETA + (a), i, j, k) * F(b, sf + (a), i, j, k))
        owned_hi(b, L_EX, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    double v = (BB(1, i, j, k - 1) - BB(1, i, j, k)) + (BB(2, i, j, k) - BB(2, i, j - 1, k));
                    F(b, se + 0, i, j, k) = 1.0 * v;
                    F(b, se + 0, i, j, k) *= F(b, S_ALPHA + 0, i, j, k);
                }
        owned_hi(b, L_EY, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    double v = (BB(0, i, j, k) - BB(0, i, j, k - 1)) + (BB(2, i - 1, j, k) - BB(2, i, j, k));
                    F(b, se + 1, i, j, k) = 1.0 * v;
                    F(b, se + 1, i, j, k) *= F(b, S_ALPHA + 1, i, j, k);
                }
        owned_hi(b, L_EZ, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    double v = (BB(0, i, j - 1, k) - BB(0, i, j, k)) + (BB(1, i, j, k) - BB(1, i - 1, j, k));
                    F(b, se + 2, i, j, k) = 1.0 * v;
                    F(b, se + 2, i, j, k) *= F(b, S_ALPHA + 2, i, j, k);
                }
#undef BB
    }
}

/* Calc_J_Edge (MercurySolver_Current.cpp:140-201): J = alpha * CurlAdjFaceToEdge(beta * B) */
static void calc_jedge(Orc *o) { adj_curl(o, S_B, S_J); }

/* calc_Jcell (MercurySolver_Current.cpp:203-259), owned cells */
static void calc_jcell(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR2
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k) {
                    double s[12] = {
                        F(b, S_J + 0, i, j, k), F(b, S_J + 0, i, j + 1, k), F(b, S_J + 0, i, j, k + 1), F(b, S_J + 0, i, j + 1, k + 1),
                        F(b, S_J + 1, i, j, k), F(b, S_J + 1, i + 1, j, k), F(b, S_J + 1, i, j, k + 1), F(b, S_J + 1, i + 1, j, k + 1),
                        F(b, S_J + 2, i, j, k), F(b, S_J + 2, i + 1, j, k), F(b, S_J + 2, i, j + 1, k), F(b, S_J + 2, i + 1, j + 1, k)};
                    double J[3] = {0, 0, 0};
                    for (int n = 0; n < 12; ++n)
                        for (int c = 0; c < 3; ++c) J[c] += F(b, S_WJ + 12 * c + n, i, j, k) * s[n];
                    for (int c = 0; c < 3; ++c) F(b, S_JC + c, i, j, k) = J[c];
                }
    }
}

/* UpdateDerivedFields_ (MercurySolver_Tools.cpp:303-321) */
static void update_derived(Orc *o) {
    calc_bcell(o);
    sync_group(o, G_BCELL);
    calc_jedge(o);
    alias_reduce_edges(o, S_J);
    singular_current(o, S_B, S_J);
    sync_group(o, G_JEDGE);
    calc_jcell(o);
    sync_group(o, G_JCELL);
    calc_pv(o);
    calc_uplus(o);
}

/* Compute_Timestep (MercurySolver_Timestep.cpp:8-141) */
static double compute_timestep(Orc *o) {
    const OrcPhys *P = &o->P;
    double dt_local = 1e100;
    for (int s = 0; s < n_species(P); ++s)
        for (int ib = 0; ib < o->nb; ++ib) {
            OBlock *b = &o->b[ib];
            int su = SP_U[s];
            /* the minimum is exact, so the reduction order does not matter */
#pragma omp parallel for collapse(2) schedule(static) reduction(min : dt_local)
            for (int i = 0; i < b->m[0]; ++i)
                for (int j = 0; j < b->m[1]; ++j)
                    for (int k = 0; k < b->m[2]; ++k) {
                        double V = fabs(F(b, S_JAC, i, j, k));
                        if (V <= 0.0) continue;
                        double rho = dmax(F(b, su, i, j, k), 1e-2);
                        double ux = F(b, su + 1, i, j, k) / rho, uy = F(b, su + 2, i, j, k) / rho, uz = F(b, su + 3, i, j, k) / rho;
                        const double E = F(b, su + 4, i, j, k);
                        const double ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
                        double p = (P->gamma - 1.0) * (E - ke);
                        double cs2 = P->gamma * p / rho;
                        double Bx = F(b, S_BC, i, j, k), By = F(b, S_BC + 1, i, j, k), Bz = F(b, S_BC + 2, i, j, k);
                        double B2 = Bx * Bx + By * By + Bz * Bz;
                        double denom = 0.0;
                        for (int a = 0; a < 3; ++a)
                            for (int sd = 0; sd < 2; ++sd) {
                                int fi = i + (a == 0 && sd), fj = j + (a == 1 && sd), fk = k + (a == 2 && sd);
                                double ax = F(b, S_JD + 3 * a, fi, fj, fk), ay = F(b, S_JD + 3 * a + 1, fi, fj, fk), az = F(b, S_JD + 3 * a + 2, fi, fj, fk);
                                double A = sqrt(ax * ax + ay * ay + az * az);
                                double term = 0.0;
                                if (!(A <= 0.0)) {
                                    double nx = ax / A, ny = ay / A, nz = az / A;
                                    if (!sd) { nx = -nx; ny = -ny; nz = -nz; }
                                    double un = ux * nx + uy * ny + uz * nz;
                                    double Bn = Bx * nx + By * ny + Bz * nz;
                                    double vA2 = P->inver_MA2 * B2 / rho;
                                    double vAn2 = P->inver_MA2 * (Bn * Bn) / rho;
                                    double t = cs2 + vA2;
                                    double disc = t * t - 4.0 * cs2 * vAn2;
                                    if (disc < 0.0) disc = 0.0;
                                    double cf = sqrt(0.5 * (t + sqrt(disc)));
                                    term = (fabs(un) + cf) * A;
                                }
                                denom += term;
                            }
                        if (denom > 0.0) {
                            double dtc = P->cfl * V / denom;
                            dt_local = dmin(dt_local, dtc);
                        }
                    }
        }
    return dt_local;
}

/* Scheme_U_ (MercurySolver_FluidRHS.cpp:3-202): Rusanov face fluxes on owned
 * faces of each axis, then RHS -= dF/Jac on owned cells.  RHS starts at zero
 * (ZeroRHS_, MercurySolver_Advance.cpp:4-71). */
static void scheme_u(Orc *o) {
    const double gm = o->P.gamma;
    for (int s = 0; s < n_species(&o->P); ++s)
        for (int ib = 0; ib < o->nb; ++ib) {
            OBlock *b = &o->b[ib];
            int su = SP_U[s], sp = SP_PV[s], sr = SP_R[s];
            for (int a = 0; a < 3; ++a) {
                int hi[3];
                owned_hi(b, L_FX + a, hi);
                PAR2
                for (int i = 0; i < hi[0]; ++i)
                    for (int j = 0; j < hi[1]; ++j)
                        for (int k = 0; k < hi[2]; ++k) {
                            const double k1 = F(b, S_JD + 3 * a, i, j, k), k2 = F(b, S_JD + 3 * a + 1, i, j, k), k3 = F(b, S_JD + 3 * a + 2, i, j, k);
                            int iL = i - (a == 0), jL = j - (a == 1), kL = k - (a == 2);
                            double W[2][5];
                            for (int side = 0; side < 2; ++side) {
                                int ii = side ? i : iL, jj = side ? j : jL, kk = side ? k : kL;
                                W[side][0] = F(b, su, ii, jj, kk);
                                for (int c = 0; c < 4; ++c) W[side][c + 1] = F(b, sp + c, ii, jj, kk);
                            }
                            double Uc[2][5], Fl[2][5], rad[2];
                            for (int side = 0; side < 2; ++side) {
                                const double rho = W[side][0], u = W[side][1], v = W[side][2], w = W[side][3], p = W[side][4];
                                Uc[side][0] = rho;
                                Uc[side][1] = rho * u;
                                Uc[side][2] = rho * v;
                                Uc[side][3] = rho * w;
                                Uc[side][4] = p / (gm - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
                                const double uvw = k1 * u + k2 * v + k3 * w;
                                const double kk2 = (k1 * k1 + k2 * k2 + k3 * k3);
                                const double c = sqrt(dmax(0.0, gm * p / dmax(rho, 1e-30)));
                                rad[side] = fabs(uvw) + c * sqrt(kk2);
                                const double rhoe = p / (gm - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
                                Fl[side][0] = rho * uvw;
                                Fl[side][1] = rho * uvw * u + k1 * p;
                                Fl[side][2] = rho * uvw * v + k2 * p;
                                Fl[side][3] = rho * uvw * w + k3 * p;
                                Fl[side][4] = uvw * (rhoe + p);
                            }
                            const double r = dmax(rad[0], rad[1]);
                            for (int m = 0; m < 5; ++m)
                                F(b, S_F + 5 * a + m, i, j, k) = 0.5 * (Fl[0][m] + Fl[1][m]) - 0.5 * r * (Uc[1][m] - Uc[0][m]);
                        }
            }
            PAR2
            for (int i = 0; i < b->m[0]; ++i)
                for (int j = 0; j < b->m[1]; ++j)
                    for (int k = 0; k < b->m[2]; ++k) {
                        const double invJ = 1.0 / F(b, S_JAC, i, j, k);
                        for (int m = 0; m < 5; ++m) {
                            double r = 0.0;
                            r -= (F(b, S_F + 0 + m, i + 1, j, k) - F(b, S_F + 0 + m, i, j, k)) * invJ;
                            r -= (F(b, S_F + 5 + m, i, j + 1, k) - F(b, S_F + 5 + m, i, j, k)) * invJ;
                            r -= (F(b, S_F + 10 + m, i, j, k + 1) - F(b, S_F + 10 + m, i, j, k)) * invJ;
                            F(b, sr + m, i, j, k) = r;
                        }
                    }
        }
}

/* electron pressure stand-in of the ambipolar terms: the sum of the species
 * pressures at a cell (p_H + p_Na in the reference, + p_He with He++) */
static double pe_sum(const OrcPhys *P, const OBlock *b, int i, int j, int k) {
    double pe = F(b, S_PVH + 3, i, j, k) + F(b, S_PVN + 3, i, j, k);
    if (n_species(P) > 2) pe = pe + F(b, S_PVHE + 3, i, j, k);
    return pe;
}

/* AddSourceToRHS_Fluid (MercurySolver_FluidRHS.cpp:206-527) */
static void add_sources(Orc *o) {
    const OrcPhys *P = &o->P;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR2
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k) {
                    const double upx = F(b, S_UP, i, j, k), upy = F(b, S_UP + 1, i, j, k), upz = F(b, S_UP + 2, i, j, k);
                    const double Bx = F(b, S_BC, i, j, k), By = F(b, S_BC + 1, i, j, k), Bz = F(b, S_BC + 2, i, j, k);
                    const double Jx = F(b, S_JC, i, j, k), Jy = F(b, S_JC + 1, i, j, k), Jz = F(b, S_JC + 2, i, j, k);
                    const double sjbx = Jy * Bz - Jz * By, sjby = Jz * Bx - Jx * Bz, sjbz = Jx * By - Jy * Bx;
                    const int ns = n_species(P);
                    NumInfo num = num_at(P, b, IX(b, i, j, k));
                    double dpex = 0.0, dpey = 0.0, dpez = 0.0;
                    if (P->ambipolar_enabled != 0.0) {
                        double V = fabs(F(b, S_JAC, i, j, k));
                        double d[9];
                        if (V <= 0.0) {
                            for (int q = 0; q < 9; ++q) d[q] = 0.0;
                        } else {
                            for (int c = 0; c < 3; ++c) {
                                d[c] = 0.5 * (F(b, S_JD + c, i - 1, j, k) + F(b, S_JD + c, i, j, k)) / V;
                                d[3 + c] = 0.5 * (F(b, S_JD + 3 + c, i, j - 1, k) + F(b, S_JD + 3 + c, i, j, k)) / V;
                                d[6 + c] = 0.5 * (F(b, S_JD + 6 + c, i, j, k - 1) + F(b, S_JD + 6 + c, i, j, k)) / V;
                            }
                        }
#define PE(a1, a2, a3) pe_sum(P, b, a1, a2, a3)
                        const double pip = PE(i + 1, j, k), pim = PE(i - 1, j, k), pjp = PE(i, j + 1, k), pjm = PE(i, j - 1, k), pkp = PE(i, j, k + 1), pkm = PE(i, j, k - 1);
#undef PE
                        dpex = 0.5 * (d[0] * (pip - pim) + d[3] * (pjp - pjm) + d[6] * (pkp - pkm));
                        dpey = 0.5 * (d[1] * (pip - pim) + d[4] * (pjp - pjm) + d[7] * (pkp - pkm));
                        dpez = 0.5 * (d[2] * (pip - pim) + d[5] * (pjp - pjm) + d[8] * (pkp - pkm));
                    }
                    const double sss = F(b, S_PHOTO, i, j, k);
                    const double x = F(b, S_CC, i, j, k), y = F(b, S_CC + 1, i, j, k), z = F(b, S_CC + 2, i, j, k);
                    const double rho_cyl = sqrt(y * y + z * z);
                    const int in_shadow = (x < 0.0 && rho_cyl < 1.0);
                    const double vst = in_shadow ? P->nu_shadow : P->nu_lit;
                    for (int s = 0; s < ns; ++s) {
                        int su = SP_U[s], sp = SP_PV[s], sr = SP_R[s];
                        const double rho = dmax(F(b, su, i, j, k), 0.0);
                        const double u = F(b, sp, i, j, k), v = F(b, sp + 1, i, j, k), w = F(b, sp + 2, i, j, k);
                        const double u2 = u * u + v * v + w * w;
                        const double subx = (v - upy) * Bz - (w - upz) * By;
                        const double suby = (w - upz) * Bx - (u - upx) * Bz;
                        const double subz = (u - upx) * By - (v - upy) * Bx;
                        const double subu = subx * u + suby * v + subz * w;
                        const double sjbu = sjbx * u + sjby * v + sjbz * w;
                        const double dpeu = dpex * u + dpey * v + dpez * w;
                        const double Ts0 = F(b, sp + 4, i, j, k) * P->T_ref;
                        const double Ms = species_mass(P, s);
                        const double Mo = P->M_Na; /* H+ pairs with (M_H, M_Na), Na+ with (M_Na, M_Na), He++ with (M_He, M_Na) */
                        const double b1 = vst * (P->Tn0 - Ts0) * (Ms / (Ms + Mo));
                        const double b2 = (Mo / (Ms + Mo)) * vst;
                        /* number density in the thermal exchange, charge density in the
                         * induction and Hall forces (equal for the singly charged ions) */
                        const double nsp = num.n[s], qs = num.q[s];
                        const double chi = num.w[s];
                        F(b, sr + 1, i, j, k) += P->c_ind * qs * subx + P->c_hall * chi * sjbx;
                        F(b, sr + 2, i, j, k) += P->c_ind * qs * suby + P->c_hall * chi * sjby;
                        F(b, sr + 3, i, j, k) += P->c_ind * qs * subz + P->c_hall * chi * sjbz;
                        F(b, sr + 4, i, j, k) += P->c_ind * qs * subu + P->c_hall * chi * sjbu;
                        if (P->ambipolar_enabled != 0.0) {
                            F(b, sr + 1, i, j, k) -= chi * dpex;
                            F(b, sr + 2, i, j, k) -= chi * dpey;
                            F(b, sr + 3, i, j, k) -= chi * dpez;
                            F(b, sr + 4, i, j, k) -= chi * dpeu;
                        }
                        if (s == 1) {
                            F(b, sr + 0, i, j, k) += P->a1_Na * sss;
                            F(b, sr + 4, i, j, k) += P->a6 * sss * P->Tn0;
                        }
                        F(b, sr + 1, i, j, k) -= P->a4 * rho * u * vst;
                        F(b, sr + 2, i, j, k) -= P->a4 * rho * v * vst;
                        F(b, sr + 3, i, j, k) -= P->a4 * rho * w * vst;
                        F(b, sr + 4, i, j, k) -= P->a4 * rho * u2 * vst;
                        F(b, sr + 4, i, j, k) += P->a5 * nsp * b1;
                        F(b, sr + 4, i, j, k) += P->a4 * rho * u2 * b2;
                    }
                }
    }
}

/* ------------------------------------------------------------------------- */
/* singular edges (valence 3 or more than 4: the corner lines of a cubed      */
/* sphere)                                                                    */
/* ------------------------------------------------------------------------- */
/* SingularEdgeRegistry::build (src/7_metric/SingularEdgeRegistry.cpp:78-268):
 * an edge class of three or more members whose aliases touch 3 or more than 4
 * distinct owner faces of positive area.  Per entry: the cells around every
 * alias (owned, |Jac| > 1e-30, weight = |Jac| / sum), the owner faces around
 * every alias (with the alias's orientation and the face's incidence on it),
 * the mean alias length, the orientation-signed mean of the aliases' dr, and
 * the dual area: the polygon through the cells' centres ordered by angle about
 * the edge (half the norm of the summed cross products), else sum|Jac| / length;
 * inverse_hodge = length / dual area, which also replaces alpha on every owned
 * alias. */
typedef struct {
    int ob, oax, osign;
    long oq;
    double cv[3], ih;
    int c0, c1, f0, f1;
} OSing;
typedef struct OSingRegT {
    int nsing, ncell, nface;
    OSing *e;
    int *cb, *cax, *cor;
    long *cq;
    double *cw;
    int *fb, *fax;
    long *fq;
    double *fcoef;
} OSingReg;

static void ijk_of(const OBlock *b, long q, int p[3]) {
    p[2] = (int)(q % b->N[2]) - b->g;
    p[1] = (int)((q / b->N[2]) % b->N[1]) - b->g;
    p[0] = (int)(q / ((long)b->N[1] * b->N[2])) - b->g;
}
static int in_owned(const OBlock *b, int L, const int p[3]) {
    int hi[3];
    owned_hi(b, L, hi);
    for (int a = 0; a < 3; ++a)
        if (p[a] < 0 || p[a] >= hi[a]) return 0;
    return 1;
}
/* incidence of a face on an edge of the same block (AssembleSingularEdgeCurrent_,
 * Z4_Mercury/src/physics/MercurySolver_Current.cpp:64-110) */
static int incidence(int ea, const int e[3], int fa, const int f[3]) {
    if (ea == 0) {
        if (fa == 1 && f[0] == e[0] && f[1] == e[1]) return f[2] == e[2] - 1 ? 1 : (f[2] == e[2] ? -1 : 0);
        if (fa == 2 && f[0] == e[0] && f[2] == e[2]) return f[1] == e[1] ? 1 : (f[1] == e[1] - 1 ? -1 : 0);
    } else if (ea == 1) {
        if (fa == 0 && f[0] == e[0] && f[1] == e[1]) return f[2] == e[2] ? 1 : (f[2] == e[2] - 1 ? -1 : 0);
        if (fa == 2 && f[1] == e[1] && f[2] == e[2]) return f[0] == e[0] - 1 ? 1 : (f[0] == e[0] ? -1 : 0);
    } else {
        if (fa == 0 && f[0] == e[0] && f[2] == e[2]) return f[1] == e[1] - 1 ? 1 : (f[1] == e[1] ? -1 : 0);
        if (fa == 1 && f[1] == e[1] && f[2] == e[2]) return f[0] == e[0] ? 1 : (f[0] == e[0] - 1 ? -1 : 0);
    }
    return 0;
}
static int face_is_alias(const Orc *o, int b, int fa, long fq) {
    for (int L = L_FX; L <= L_FZ; ++L)
        for (int c = 0; c < o->ncls[L]; ++c)
            for (long t = o->cls_start[L][c] + 1; t < o->cls_start[L][c + 1]; ++t)
                if (o->cls_b[L][t] == b && o->cls_q[L][t] == fq && o->cls_ax[L][t] == fa) return 1;
    return 0;
}
typedef struct { double ang, p[3]; } OAng;
static int ang_cmp(const void *x, const void *y) {
    const double a = ((const OAng *)x)->ang, b = ((const OAng *)y)->ang;
    return (a > b) - (a < b);
}
static int build_singular(Orc *o) {
    OSingReg *S = (OSingReg *)calloc(1, sizeof(OSingReg));
    o->singp = S;
    const int L = L_EX;
    for (int c = 0; c < o->ncls[L]; ++c) {
        const long s0 = o->cls_start[L][c], s1 = o->cls_start[L][c + 1];
        if (s1 - s0 < 3) continue;
        int nc = 0, nf = 0, nlen = 0;
        int cbuf[64][4], fbuf[64][5];
        double cw[64], len = 0.0, dr[3] = {0, 0, 0}, vsum = 0.0, fsum = 0.0;
        int cax[64], cor[64], fcoef_sign[64];
        double cen[64][3];
        int ok = 1;
        for (long t = s0; t < s1 && ok; ++t) {
            const int b = o->cls_b[L][t], ax = o->cls_ax[L][t], orient = o->cls_sg[L][t];
            OBlock *bk = &o->b[b];
            int e[3];
            ijk_of(bk, o->cls_q[L][t], e);
            if (in_owned(bk, L_EX + ax, e)) {
                const double Lg = F(bk, S_DL + ax, e[0], e[1], e[2]);
                if (isfinite(Lg) && Lg > 0.0) len += Lg, ++nlen;
                for (int q = 0; q < 3; ++q) dr[q] += orient * F(bk, S_DR + 3 * ax + q, e[0], e[1], e[2]);
            }
            for (int s0c = -1; s0c <= 0; ++s0c)
                for (int s1c = -1; s1c <= 0; ++s1c) {
                    int p[3] = {e[0], e[1], e[2]};
                    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
                    p[t1] += s0c;
                    p[t2] += s1c;
                    int dup = 0;
                    for (int u = 0; u < nc && !dup; ++u) dup = cbuf[u][0] == b && cbuf[u][1] == p[0] && cbuf[u][2] == p[1] && cbuf[u][3] == p[2];
                    if (dup || !in_owned(bk, L_CELL, p)) continue;
                    const double v = fabs(F(bk, S_JAC, p[0], p[1], p[2]));
                    if (!(isfinite(v) && v > 1.e-30)) continue;
                    if (nc == 64) { ok = 0; break; }
                    cbuf[nc][0] = b, cbuf[nc][1] = p[0], cbuf[nc][2] = p[1], cbuf[nc][3] = p[2];
                    cw[nc] = v, cax[nc] = ax, cor[nc] = orient;
                    for (int q = 0; q < 3; ++q) cen[nc][q] = F(bk, S_CC + q, p[0], p[1], p[2]);
                    vsum += v;
                    ++nc;
                }
            /* the four faces around the edge (incident_faces, SingularEdgeRegistry.cpp:43-68) */
            int fl[4][4];
            if (ax == 0) {
                int t0[4][4] = {{1, e[0], e[1], e[2]}, {1, e[0], e[1], e[2] - 1}, {2, e[0], e[1], e[2]}, {2, e[0], e[1] - 1, e[2]}};
                memcpy(fl, t0, sizeof fl);
            } else if (ax == 1) {
                int t0[4][4] = {{0, e[0], e[1], e[2]}, {0, e[0], e[1], e[2] - 1}, {2, e[0], e[1], e[2]}, {2, e[0] - 1, e[1], e[2]}};
                memcpy(fl, t0, sizeof fl);
            } else {
                int t0[4][4] = {{0, e[0], e[1], e[2]}, {0, e[0], e[1] - 1, e[2]}, {1, e[0], e[1], e[2]}, {1, e[0] - 1, e[1], e[2]}};
                memcpy(fl, t0, sizeof fl);
            }
            for (int u = 0; u < 4; ++u) {
                const int fa = fl[u][0], fp[3] = {fl[u][1], fl[u][2], fl[u][3]};
                int dup = 0;
                for (int w = 0; w < nf && !dup; ++w)
                    dup = fbuf[w][0] == b && fbuf[w][1] == fa && fbuf[w][2] == fp[0] && fbuf[w][3] == fp[1] && fbuf[w][4] == fp[2];
                if (dup || !in_owned(bk, L_FX + fa, fp)) continue;
                /* only owner faces (Topology::is_owner): no later member of a face class */
                if (face_is_alias(o, b, fa, IX(bk, fp[0], fp[1], fp[2]))) continue;
                const double av = F(bk, S_AREA + fa, fp[0], fp[1], fp[2]);
                if (!(isfinite(av) && av > 1.e-30)) continue;
                if (nf == 64) { ok = 0; break; }
                fbuf[nf][0] = b, fbuf[nf][1] = fa, fbuf[nf][2] = fp[0], fbuf[nf][3] = fp[1], fbuf[nf][4] = fp[2];
                fcoef_sign[nf] = orient * incidence(ax, e, fa, fp);
                fsum += av;
                ++nf;
            }
        }
        if (!ok) return 0;
        if (!(nf == 3 || nf > 4)) continue;
        OSing E;
        E.ob = o->cls_b[L][s0], E.oq = o->cls_q[L][s0], E.oax = o->cls_ax[L][s0], E.osign = o->cls_sg[L][s0];
        const double plen = nlen ? len / nlen : 0.0;
        double dual = 0.0;
        if (nc >= 3) {
            double ctr[3] = {0, 0, 0}, tg[3] = {dr[0], dr[1], dr[2]};
            for (int u = 0; u < nc; ++u)
                for (int q = 0; q < 3; ++q) ctr[q] += cen[u][q];
            for (int q = 0; q < 3; ++q) ctr[q] /= nc;
            const double tn = sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
            if (tn > 0.0)
                for (int q = 0; q < 3; ++q) tg[q] /= tn;
            double ref[3] = {1, 0, 0};
            if (!(fabs(tg[0]) < 0.9)) ref[0] = 0, ref[1] = 1;
            double uu[3] = {tg[1] * ref[2] - tg[2] * ref[1], tg[2] * ref[0] - tg[0] * ref[2], tg[0] * ref[1] - tg[1] * ref[0]};
            const double un = sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
            if (un > 0.0)
                for (int q = 0; q < 3; ++q) uu[q] /= un;
            const double vv[3] = {tg[1] * uu[2] - tg[2] * uu[1], tg[2] * uu[0] - tg[0] * uu[2], tg[0] * uu[1] - tg[1] * uu[0]};
            OAng pts[64];
            for (int u = 0; u < nc; ++u) {
                double du = 0.0, dv = 0.0;
                for (int q = 0; q < 3; ++q) {
                    du += (cen[u][q] - ctr[q]) * uu[q];
                    dv += (cen[u][q] - ctr[q]) * vv[q];
                }
                pts[u].ang = atan2(dv, du);
                for (int q = 0; q < 3; ++q) pts[u].p[q] = cen[u][q];
            }
            qsort(pts, nc, sizeof(OAng), ang_cmp);
            double av[3] = {0, 0, 0};
            for (int u = 0; u < nc; ++u) {
                const double *a = pts[u].p, *bb = pts[(u + 1) % nc].p;
                av[0] += a[1] * bb[2] - a[2] * bb[1];
                av[1] += a[2] * bb[0] - a[0] * bb[2];
                av[2] += a[0] * bb[1] - a[1] * bb[0];
            }
            dual = 0.5 * sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
        }
        if (!(dual > 0.0)) dual = plen > 0.0 ? vsum / plen : 0.0;
        E.ih = dual > 0.0 ? plen / dual : 0.0;
        for (int q = 0; q < 3; ++q) E.cv[q] = nlen > 0 ? dr[q] / nlen : 0.0;
        /* validate_or_abort (SingularEdgeRegistry.cpp:270-282) */
        if (!(plen > 0.0) || !(dual > 0.0) || !isfinite(E.ih) || nc < 1) return 0;
        (void)fsum;
        E.c0 = S->ncell, E.c1 = S->ncell + nc, E.f0 = S->nface, E.f1 = S->nface + nf;
        S->e = (OSing *)realloc(S->e, sizeof(OSing) * (S->nsing + 1));
        S->e[S->nsing++] = E;
        S->cb = realloc(S->cb, sizeof(int) * (S->ncell + nc)), S->cax = realloc(S->cax, sizeof(int) * (S->ncell + nc));
        S->cor = realloc(S->cor, sizeof(int) * (S->ncell + nc)), S->cq = realloc(S->cq, sizeof(long) * (S->ncell + nc));
        S->cw = realloc(S->cw, sizeof(double) * (S->ncell + nc));
        for (int u = 0; u < nc; ++u) {
            const int n = S->ncell + u;
            S->cb[n] = cbuf[u][0], S->cq[n] = IX(&o->b[cbuf[u][0]], cbuf[u][1], cbuf[u][2], cbuf[u][3]);
            S->cw[n] = vsum > 0.0 ? cw[u] / vsum : 0.0, S->cax[n] = cax[u], S->cor[n] = cor[u];
        }
        S->ncell += nc;
        S->fb = realloc(S->fb, sizeof(int) * (S->nface + nf)), S->fax = realloc(S->fax, sizeof(int) * (S->nface + nf));
        S->fq = realloc(S->fq, sizeof(long) * (S->nface + nf)), S->fcoef = realloc(S->fcoef, sizeof(double) * (S->nface + nf));
        for (int w = 0; w < nf; ++w) {
            const int n = S->nface + w;
            S->fb[n] = fbuf[w][0], S->fax[n] = fbuf[w][1], S->fq[n] = IX(&o->b[fbuf[w][0]], fbuf[w][2], fbuf[w][3], fbuf[w][4]);
            S->fcoef[n] = E.ih * fcoef_sign[w];
        }
        S->nface += nf;
        /* the corrected inverse Hodge on every owned alias */
        for (long t = s0; t < s1; ++t) {
            OBlock *bk = &o->b[o->cls_b[L][t]];
            int e[3];
            ijk_of(bk, o->cls_q[L][t], e);
            if (in_owned(bk, L_EX + o->cls_ax[L][t], e)) bk->f[S_ALPHA + o->cls_ax[L][t]][o->cls_q[L][t]] = E.ih;
        }
    }
    return 1;
}

static double hall_taper_at(const OrcPhys *P, double x, double y, double z) {
    const double r = sqrt(x * x + y * y + z * z);
    const double width = P->taper_rmax - P->taper_rmin;
    if (width <= 0.0) return (r >= P->taper_rmin) ? 1.0 : 0.0;
    const double mid = 0.5 * (P->taper_rmin + P->taper_rmax);
    return 0.5 * (1.0 + tanh(2.0 * (r - mid) / width));
}
/* AssembleSingularEdgeEMF_NonHall_ (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Ideal.cpp:57-130)
 * (hall = 0) and AssembleSingularEdgeEMF_HallExplicit_ (MercurySolver_MagExplicit_Hall.cpp:32-56)
 * (hall = 1): owner = owner sign * sum over the cells of weight * the cell's EMF
 * along the mean edge vector (assemble_cell_field_to_local_owners,
 * SingularEdgeRegistry.cpp:288-310).  The electron pressure and the number
 * limiter take the H and Na species only, as there. */
static void singular_emf(Orc *o, int hall, int slot0) {
    const OrcPhys *P = &o->P;
    const OSingReg *S = o->singp;
    if (!S) return;
    for (int n = 0; n < S->nsing; ++n) {
        const OSing *E = &S->e[n];
        double sum = 0.0;
        for (int t = E->c0; t < E->c1; ++t) {
            const OBlock *b = &o->b[S->cb[t]];
            const long q = S->cq[t];
            const double *cv = E->cv;
            double emf;
            const double b0 = b->f[S_BC][q], b1 = b->f[S_BC + 1][q], b2 = b->f[S_BC + 2][q];
            const double j0 = b->f[S_JC][q], j1 = b->f[S_JC + 1][q], j2 = b->f[S_JC + 2][q];
            const double rho2[NSMAX] = {b->f[S_UH][q], b->f[S_UN][q], 0.0};
            if (!hall) {
                const double u0 = b->f[S_UP][q], u1 = b->f[S_UP + 1][q], u2 = b->f[S_UP + 2][q];
                const double ex = -(u1 * b2 - u2 * b1), ey = -(u2 * b0 - u0 * b2), ez = -(u0 * b1 - u1 * b0);
                emf = ex * cv[0] + ey * cv[1] + ez * cv[2];
                if (P->ambipolar_enabled != 0.0) {
                    const long d = S->cax[t] == 0 ? (long)b->N[1] * b->N[2] : (S->cax[t] == 1 ? (long)b->N[2] : 1L);
                    const double dp = 0.5 * ((b->f[S_PVH + 3][q + d] + b->f[S_PVN + 3][q + d]) -
                                             (b->f[S_PVH + 3][q - d] + b->f[S_PVN + 3][q - d]));
                    const NumInfo num = num_limiter(P, rho2, 2);
                    emf -= P->pressure_coef * S->cor[t] * dp / num.ne_eff;
                }
                const double jedge = j0 * cv[0] + j1 * cv[1] + j2 * cv[2];
                const double jmag = sqrt(j0 * j0 + j1 * j1 + j2 * j2);
                if (P->arti_eta_max > 0.0) {
                    double qq = (jmag - P->J_start) / dmax(P->J_on - P->J_start, 1.e-30);
                    qq = dmax(0.0, dmin(1.0, qq));
                    emf += P->arti_eta_max * qq * qq * (3.0 - 2.0 * qq) * jedge;
                }
                if (P->local_enabled != 0.0 && P->local_eta_max > 0.0) {
                    const double dx = b->f[S_CC][q] - P->local_center[0], dy = b->f[S_CC + 1][q] - P->local_center[1],
                                 dz = b->f[S_CC + 2][q] - P->local_center[2];
                    const double rr = sqrt(dx * dx + dy * dy + dz * dz);
                    if (P->local_r_cutoff <= 0.0 || rr < P->local_r_cutoff) emf += P->local_eta_max * exp(-rr / P->local_r_decay) * jedge;
                }
            } else {
                const NumInfo num = num_limiter(P, rho2, 2);
                const double alpha = P->hall_coef / num.ne_eff;
                const double ex = alpha * (j1 * b2 - j2 * b1), ey = alpha * (j2 * b0 - j0 * b2), ez = alpha * (j0 * b1 - j1 * b0);
                const double taper = hall_taper_at(P, b->f[S_CC][q], b->f[S_CC + 1][q], b->f[S_CC + 2][q]);
                emf = taper * (ex * cv[0] + ey * cv[1] + ez * cv[2]);
            }
            sum += S->cw[t] * emf;
        }
        o->b[E->ob].f[slot0 + E->oax][E->oq] = E->osign * sum;
    }
}
/* AssembleSingularEdgeCurrent_ (Z4_Mercury/src/physics/MercurySolver_Current.cpp:59-138):
 * owner = owner sign * sum over the owner faces of inverse_hodge * alias
 * orientation * incidence * beta * B (assemble_face_triplet_to_local_owners,
 * SingularEdgeRegistry.cpp:312-330) */
static void singular_current(Orc *o, int sB, int sJ) {
    const OSingReg *S = o->singp;
    if (!S) return;
    for (int n = 0; n < S->nsing; ++n) {
        const OSing *E = &S->e[n];
        double sum = 0.0;
        for (int t = E->f0; t < E->f1; ++t) {
            const OBlock *b = &o->b[S->fb[t]];
            sum += S->fcoef[t] * b->f[S_BETA + S->fax[t]][S->fq[t]] * b->f[sB + S->fax[t]][S->fq[t]];
        }
        o->b[E->ob].f[sJ + E->oax][E->oq] = E->osign * sum;
    }
}

/* helpers of AssembleOneDirectionEMF_ (MercurySolver_MagExplicit_Ideal.cpp:433-594) */
static inline double u_contra(const OBlock *b, int axis, int i, int j, int k) {
    int ip = i + (axis == 0), jp = j + (axis == 1), kp = k + (axis == 2);
    const int s0 = S_JD + 3 * axis;
    const double kx = 0.5 * (F(b, s0, i, j, k) + F(b, s0, ip, jp, kp));
    const double ky = 0.5 * (F(b, s0 + 1, i, j, k) + F(b, s0 + 1, ip, jp, kp));
    const double kz = 0.5 * (F(b, s0 + 2, i, j, k) + F(b, s0 + 2, ip, jp, kp));
    const double inv_jac = 1.0 / (fabs(F(b, S_JAC, i, j, k)) + 1.0e-30);
    return (kx * F(b, S_UP, i, j, k) + ky * F(b, S_UP + 1, i, j, k) + kz * F(b, S_UP + 2, i, j, k)) * inv_jac;
}
static inline double beta_ind(const OBlock *b, int axis, int i, int j, int k) {
    return 0.5 * (F(b, S_B + axis, i, j, k) + F(b, S_B + axis, i + (axis == 0), j + (axis == 1), k + (axis == 2)));
}
static inline double beta_tot(const OBlock *b, int axis, int i, int j, int k) {
    int ip = i + (axis == 0), jp = j + (axis == 1), kp = k + (axis == 2);
    return 0.5 * ((F(b, S_B + axis, i, j, k) + F(b, S_BADD + axis, i, j, k)) + (F(b, S_B + axis, ip, jp, kp) + F(b, S_BADD + axis, ip, jp, kp)));
}
static inline double emf_flux(const OBlock *b, int ba, int fa, int iL, int jL, int kL, int iR, int jR, int kR, double bff) {
    const double btL = beta_tot(b, ba, iL, jL, kL), btR = beta_tot(b, ba, iR, jR, kR);
    const double biL = beta_ind(b, ba, iL, jL, kL), biR = beta_ind(b, ba, iR, jR, kR);
    const double ufL = u_contra(b, fa, iL, jL, kL), ufR = u_contra(b, fa, iR, jR, kR);
    const double ubL = u_contra(b, ba, iL, jL, kL), ubR = u_contra(b, ba, iR, jR, kR);
    const double GL = ufL * btL - ubL * bff, GR = ufR * btR - ubR * bff;
    const double radius = dmax(fabs(ufL), fabs(ufR));
    return 0.5 * (GL + GR) - 0.5 * radius * (biR - biL);
}

/* AddIdealEdgeEMF_ (MercurySolver_MagExplicit_Ideal.cpp:7-55): face candidates on
 * owned faces +-1 (clamped to the allocation +-1), then
 * AssembleEdgeEMF_FromFaceE_Ideal_ (:138-372) on owned edges. */
static void ideal_emf(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int fa = 0; fa < 3; ++fa) {
            int ohi[3], alo[3], ahi[3], lo[3], hi[3];
            owned_hi(b, L_FX + fa, ohi);
            alloc_box(b, L_FX + fa, b->g, alo, ahi);
            for (int a = 0; a < 3; ++a) {
                lo[a] = dmax(0 - 1, alo[a] + 1);
                hi[a] = dmin(ohi[a] + 1, ahi[a] - 1);
            }
            PAR2
            for (int i = lo[0]; i < hi[0]; ++i)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int k = lo[2]; k < hi[2]; ++k) {
                        int iL = i - (fa == 0), jL = j - (fa == 1), kL = k - (fa == 2);
                        const double bff = F(b, S_B + fa, i, j, k) + F(b, S_BADD + fa, i, j, k);
                        double e[3] = {0, 0, 0};
                        if (fa == 0) {
                            double Gex = emf_flux(b, 1, 0, iL, jL, kL, i, j, k, bff);
                            double Gzx = emf_flux(b, 2, 0, iL, jL, kL, i, j, k, bff);
                            e[1] = Gzx;
                            e[2] = -Gex;
                        } else if (fa == 1) {
                            double Gxe = emf_flux(b, 0, 1, iL, jL, kL, i, j, k, bff);
                            double Gze = emf_flux(b, 2, 1, iL, jL, kL, i, j, k, bff);
                            e[0] = -Gze;
                            e[2] = Gxe;
                        } else {
                            double Gxz = emf_flux(b, 0, 2, iL, jL, kL, i, j, k, bff);
                            double Gez = emf_flux(b, 1, 2, iL, jL, kL, i, j, k, bff);
                            e[0] = Gez;
                            e[1] = -Gxz;
                        }
                        for (int c = 0; c < 3; ++c) F(b, S_EF + 3 * fa + c, i, j, k) = e[c];
                    }
        }
        const double Cuct = 0.5;
        int hi[3];
#define MAXU4(ax, a0, a1, a2, b0, b1, b2, c0, c1, c2, d0, d1, d2)                                                    \
    dmax(dmax(dmax(dmax(0.0, fabs(u_contra(b, ax, a0, a1, a2))), fabs(u_contra(b, ax, b0, b1, b2))), fabs(u_contra(b, ax, c0, c1, c2))), \
         fabs(u_contra(b, ax, d0, d1, d2)))
        owned_hi(b, L_EX, hi);
        PAR2
        for (int i = 0; i < hi[0]; i++)
            for (int j = 0; j < hi[1]; j++)
                for (int k = 0; k < hi[2]; k++) {
                    const double Ec = 0.25 * (F(b, S_EF + 3, i, j, k) + F(b, S_EF + 3, i, j, k - 1) + F(b, S_EF + 6, i, j, k) + F(b, S_EF + 6, i, j - 1, k));
                    const double ae = MAXU4(1, i, j, k, i, j - 1, k, i, j, k - 1, i, j - 1, k - 1);
                    const double az = MAXU4(2, i, j, k, i, j - 1, k, i, j, k - 1, i, j - 1, k - 1);
                    const double dBz = F(b, S_B + 2, i, j, k) - F(b, S_B + 2, i, j - 1, k);
                    const double dBe = F(b, S_B + 1, i, j, k) - F(b, S_B + 1, i, j, k - 1);
                    F(b, S_E + 0, i, j, k) = Ec + Cuct * (0.5 * ae * dBz - 0.5 * az * dBe);
                }
        owned_hi(b, L_EY, hi);
        PAR2
        for (int i = 0; i < hi[0]; i++)
            for (int j = 0; j < hi[1]; j++)
                for (int k = 0; k < hi[2]; k++) {
                    const double Ec = 0.25 * (F(b, S_EF + 1, i, j, k) + F(b, S_EF + 1, i, j, k - 1) + F(b, S_EF + 7, i, j, k) + F(b, S_EF + 7, i - 1, j, k));
                    const double az = MAXU4(2, i, j, k, i - 1, j, k, i, j, k - 1, i - 1, j, k - 1);
                    const double ax = MAXU4(0, i, j, k, i - 1, j, k, i, j, k - 1, i - 1, j, k - 1);
                    const double dBx = F(b, S_B + 0, i, j, k) - F(b, S_B + 0, i, j, k - 1);
                    const double dBz = F(b, S_B + 2, i, j, k) - F(b, S_B + 2, i - 1, j, k);
                    F(b, S_E + 1, i, j, k) = Ec + Cuct * (0.5 * az * dBx - 0.5 * ax * dBz);
                }
        owned_hi(b, L_EZ, hi);
        PAR2
        for (int i = 0; i < hi[0]; i++)
            for (int j = 0; j < hi[1]; j++)
                for (int k = 0; k < hi[2]; k++) {
                    const double Ec = 0.25 * (F(b, S_EF + 2, i, j, k) + F(b, S_EF + 2, i, j - 1, k) + F(b, S_EF + 5, i, j, k) + F(b, S_EF + 5, i - 1, j, k));
                    const double ax = MAXU4(0, i, j, k, i - 1, j, k, i, j - 1, k, i - 1, j - 1, k);
                    const double ae = MAXU4(1, i, j, k, i - 1, j, k, i, j - 1, k, i - 1, j - 1, k);
                    const double dBe = F(b, S_B + 1, i, j, k) - F(b, S_B + 1, i - 1, j, k);
                    const double dBx = F(b, S_B + 0, i, j, k) - F(b, S_B + 0, i, j - 1, k);
                    F(b, S_E + 2, i, j, k) = Ec + Cuct * (0.5 * ax * dBe - 0.5 * ae * dBx);
                }
#undef MAXU4
    }
}

/* AddAmbipolarEdgeEMF_ (MercurySolver_MagExplicit_Ambipo.cpp:3-143) */
static void ambipolar_emf(Orc *o) {
    const OrcPhys *P = &o->P;
    if (P->ambipolar_enabled == 0.0) return;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
#define PEC(a1, a2, a3) pe_sum(P, b, a1, a2, a3)
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_EX + a, hi);
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) {
                        double pe[2];
                        for (int e = 0; e < 2; ++e) {
                            int ni = i + (e && a == 0), nj = j + (e && a == 1), nk = k + (e && a == 2);
                            double sum = 0.0;
                            for (int di = -1; di <= 0; ++di)
                                for (int dj = -1; dj <= 0; ++dj)
                                    for (int dk = -1; dk <= 0; ++dk) sum += PEC(ni + di, nj + dj, nk + dk);
                            pe[e] = 0.125 * sum;
                        }
                        const int ns = n_species(P);
                        double rs[NSMAX] = {0.0, 0.0, 0.0};
                        for (int d1 = -1; d1 <= 0; ++d1)
                            for (int d2 = -1; d2 <= 0; ++d2) {
                                int ii = i, jj = j, kk = k;
                                if (a == 0) { jj += d1; kk += d2; }
                                else if (a == 1) { ii += d1; kk += d2; }
                                else { ii += d1; jj += d2; }
                                for (int s = 0; s < ns; ++s) rs[s] += F(b, SP_U[s], ii, jj, kk);
                            }
                        for (int s = 0; s < ns; ++s) rs[s] = 0.25 * rs[s];
                        NumInfo n = num_limiter(P, rs, ns);
                        F(b, S_E + a, i, j, k) -= P->pressure_coef * (pe[1] - pe[0]) / n.ne_eff;
                    }
        }
#undef PEC
    }
}

/* AddArtificialResistivityToEdgeEMF_ and AddLocalArtificialResistivityToEdgeEMF_
 * (MercurySolver_MagExplicit_Resis.cpp:6-185) */
static void arti_emf(Orc *o) {
    const OrcPhys *P = &o->P;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_EX + a, hi);
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) {
                        if (P->arti_eta_max > 0.0) {
                            const double J1 = F(b, S_J + a, i, j, k);
                            const double dx = F(b, S_DR + 3 * a, i, j, k), dy = F(b, S_DR + 3 * a + 1, i, j, k), dz = F(b, S_DR + 3 * a + 2, i, j, k);
                            const double drl = sqrt(dx * dx + dy * dy + dz * dz);
                            const double Jm = fabs(J1) / dmax(drl, 1.0e-30);
                            double eta;
                            if (P->J_on <= P->J_start)
                                eta = (Jm > P->J_start) ? P->arti_eta_max : 0.0;
                            else {
                                double q = (Jm - P->J_start) / (P->J_on - P->J_start);
                                q = dmax(0.0, dmin(1.0, q));
                                eta = P->arti_eta_max * (q * q * (3.0 - 2.0 * q));
                            }
                            F(b, S_E + a, i, j, k) += eta * J1;
                        }
                    }
        }
    }
    if (!(P->local_enabled != 0.0 && P->local_eta_max > 0.0 && P->local_r_decay > 0.0)) return;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_EX + a, hi);
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) F(b, S_E + a, i, j, k) += F(b, S_LETA + a, i, j, k) * F(b, S_J + a, i, j, k);
        }
    }
}

/* BuildHallFaceEMF_Rusanov_diff_ (MercurySolver_MagExplicit_Hall.cpp:58-433), no Pole patches */
static void hall_emf(Orc *o) {
    const OrcPhys *P = &o->P;
    const double eps = 1e-14, Cwh = 0.1;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR1
        for (long q = 0; q < b->V; ++q) {
            NumInfo n = num_at(P, b, q);
            const double alpha = P->hall_coef / n.ne_eff;
            const double Jx = b->f[S_JC][q], Jy = b->f[S_JC + 1][q], Jz = b->f[S_JC + 2][q];
            const double Bx = b->f[S_BC][q], By = b->f[S_BC + 1][q], Bz = b->f[S_BC + 2][q];
            b->f[S_EHC][q] = alpha * (Jy * Bz - Jz * By);
            b->f[S_EHC + 1][q] = alpha * (Jz * Bx - Jx * Bz);
            b->f[S_EHC + 2][q] = alpha * (Jx * By - Jy * Bx);
            b->f[S_HB][q] = fabs(alpha) * sqrt(Bx * Bx + By * By + Bz * Bz);
        }
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_EX + a, hi);
            /* the four cells around the edge, in the reference's order */
            int c4[4][3];
            if (a == 0) { int t[4][3] = {{0, -1, -1}, {0, -1, 0}, {0, 0, -1}, {0, 0, 0}}; memcpy(c4, t, sizeof(t)); }
            else if (a == 1) { int t[4][3] = {{-1, 0, -1}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 0}}; memcpy(c4, t, sizeof(t)); }
            else { int t[4][3] = {{-1, -1, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, 0}}; memcpy(c4, t, sizeof(t)); }
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) {
                        double Ec[3];
                        for (int c = 0; c < 3; ++c)
                            Ec[c] = 0.25 * (F(b, S_EHC + c, i + c4[0][0], j + c4[0][1], k + c4[0][2]) + F(b, S_EHC + c, i + c4[1][0], j + c4[1][1], k + c4[1][2]) +
                                            F(b, S_EHC + c, i + c4[2][0], j + c4[2][1], k + c4[2][2]) + F(b, S_EHC + c, i + c4[3][0], j + c4[3][1], k + c4[3][2]));
                        const double Ecen = Ec[0] * F(b, S_DR + 3 * a, i, j, k) + Ec[1] * F(b, S_DR + 3 * a + 1, i, j, k) + Ec[2] * F(b, S_DR + 3 * a + 2, i, j, k);
                        const double be = dmax(dmax(F(b, S_HB, i + c4[0][0], j + c4[0][1], k + c4[0][2]), F(b, S_HB, i + c4[1][0], j + c4[1][1], k + c4[1][2])),
                                               dmax(F(b, S_HB, i + c4[2][0], j + c4[2][1], k + c4[2][2]), F(b, S_HB, i + c4[3][0], j + c4[3][1], k + c4[3][2])));
                        const double L = dmax(F(b, S_DL + a, i, j, k), eps);
                        double val;
                        if (a == 0) {
                            const double Az = 0.5 * (F(b, S_AREA + 2, i, j, k) + F(b, S_AREA + 2, i, j - 1, k));
                            const double ihe = L / dmax(Az, eps);
                            const double mue = Cwh * be * ihe * ihe;
                            const double Ae = 0.5 * (F(b, S_AREA + 1, i, j, k) + F(b, S_AREA + 1, i, j, k - 1));
                            const double ihz = L / dmax(Ae, eps);
                            const double muz = Cwh * be * ihz * ihz;
                            const double d1 = F(b, S_B + 2, i, j, k) - F(b, S_B + 2, i, j - 1, k);
                            const double d2 = F(b, S_B + 1, i, j, k) - F(b, S_B + 1, i, j, k - 1);
                            val = Ecen + 0.5 * mue * d1 - 0.5 * muz * d2;
                        } else if (a == 1) {
                            const double Ax = 0.5 * (F(b, S_AREA + 0, i, j, k) + F(b, S_AREA + 0, i, j, k - 1));
                            const double ihz = L / dmax(Ax, eps);
                            const double muz = Cwh * be * ihz * ihz;
                            const double Az = 0.5 * (F(b, S_AREA + 2, i, j, k) + F(b, S_AREA + 2, i - 1, j, k));
                            const double ihx = L / dmax(Az, eps);
                            const double mux = Cwh * be * ihx * ihx;
                            const double d1 = F(b, S_B + 0, i, j, k) - F(b, S_B + 0, i, j, k - 1);
                            const double d2 = F(b, S_B + 2, i, j, k) - F(b, S_B + 2, i - 1, j, k);
                            val = Ecen + 0.5 * muz * d1 - 0.5 * mux * d2;
                        } else {
                            const double Ae = 0.5 * (F(b, S_AREA + 1, i, j, k) + F(b, S_AREA + 1, i - 1, j, k));
                            const double ihx = L / dmax(Ae, eps);
                            const double mux = Cwh * be * ihx * ihx;
                            const double Ax = 0.5 * (F(b, S_AREA + 0, i, j, k) + F(b, S_AREA + 0, i, j - 1, k));
                            const double ihe = L / dmax(Ax, eps);
                            const double mue = Cwh * be * ihe * ihe;
                            const double d1 = F(b, S_B + 1, i, j, k) - F(b, S_B + 1, i - 1, j, k);
                            const double d2 = F(b, S_B + 0, i, j, k) - F(b, S_B + 0, i, j - 1, k);
                            val = Ecen + 0.5 * mux * d1 - 0.5 * mue * d2;
                        }
                        F(b, S_EH + a, i, j, k) = F(b, S_TAPER + a, i, j, k) * val;
                    }
        }
    }
}

/* ApplyStationaryWallNonResistiveEMF_ (MercurySolver_MagneticRHS.cpp:5-89):
 * tangential E on the owned wall edges is zeroed. */
static void wall_zero_e(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int f = 0; f < 6; ++f) {
            if (b->bc[f] != BC_WALL) continue;
            int nax = f / 2;
            for (int a = 0; a < 3; ++a) {
                if (a == nax) continue;
                int hi[3], lo[3] = {0, 0, 0};
                owned_hi(b, L_EX + a, hi);
                if (f & 1) lo[nax] = b->m[nax]; else hi[nax] = 1;
                if (f & 1) hi[nax] = b->m[nax] + 1;
                for (int i = lo[0]; i < hi[0]; ++i)
                    for (int j = lo[1]; j < hi[1]; ++j)
                        for (int k = lo[2]; k < hi[2]; ++k) F(b, S_E + a, i, j, k) = 0.0;
            }
        }
    }
}

/* AssembleRHS_Induction_CT_ (MercurySolver_MagneticRHS.cpp:90-171) with
 * CTOperators::CurlEdgeToFace(-1) (CTOperators.cpp:6-57) */
static void induction_ct(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int a = 0; a < 3; ++a) memset(b->f[S_E + a], 0, sizeof(double) * b->V);
    }
    ideal_emf(o);
    ambipolar_emf(o);
    arti_emf(o);
    alias_reduce_edges(o, S_E);
    singular_emf(o, 0, S_E);
    hall_emf(o);
    singular_emf(o, 1, S_EH);
    sync_group(o, G_EHALL);
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_EX + a, hi);
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) F(b, S_E + a, i, j, k) += F(b, S_EH + a, i, j, k);
        }
    }
    wall_zero_e(o);
    sync_group(o, G_EEDGE);
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        int hi[3];
        owned_hi(b, L_FX, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    const double c = (F(b, S_E + 1, i, j, k) - F(b, S_E + 1, i, j, k + 1)) + (F(b, S_E + 2, i, j + 1, k) - F(b, S_E + 2, i, j, k));
                    F(b, S_RHSB + 0, i, j, k) = -1.0 * c;
                }
        owned_hi(b, L_FY, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    const double c = (F(b, S_E + 0, i, j, k + 1) - F(b, S_E + 0, i, j, k)) + (F(b, S_E + 2, i, j, k) - F(b, S_E + 2, i + 1, j, k));
                    F(b, S_RHSB + 1, i, j, k) = -1.0 * c;
                }
        owned_hi(b, L_FZ, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) {
                    const double c = (F(b, S_E + 0, i, j, k) - F(b, S_E + 0, i, j + 1, k)) + (F(b, S_E + 1, i + 1, j, k) - F(b, S_E + 1, i, j, k));
                    F(b, S_RHSB + 2, i, j, k) = -1.0 * c;
                }
    }
}

/* ApplyUpdate_Euler_ (MercurySolver_Advance.cpp:73-147), owned DOFs */
static void apply_update(Orc *o, double dt) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR2
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k)
                    for (int m = 0; m < 5; ++m)
                        for (int s = 0; s < n_species(&o->P); ++s) F(b, SP_U[s] + m, i, j, k) += dt * F(b, SP_R[s] + m, i, j, k);
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_FX + a, hi);
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) F(b, S_B + a, i, j, k) += dt * F(b, S_RHSB + a, i, j, k);
        }
    }
}

/* SSP-RK2 second-stage average U <- 0.5*(U0 + U), owned DOFs (no reference
 * counterpart: the reference integrator is forward Euler) */
static void rk2_average(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        PAR2
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k)
                    for (int m = 0; m < 5; ++m)
                        for (int s = 0; s < n_species(&o->P); ++s)
                            F(b, SP_U[s] + m, i, j, k) = 0.5 * (F(b, SP_U0[s] + m, i, j, k) + F(b, SP_U[s] + m, i, j, k));
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_FX + a, hi);
            PAR2
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) F(b, S_B + a, i, j, k) = 0.5 * (F(b, S_B0 + a, i, j, k) + F(b, S_B + a, i, j, k));
        }
    }
}

static void save_state(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int m = 0; m < 5; ++m)
            for (int s = 0; s < n_species(&o->P); ++s) memcpy(b->f[SP_U0[s] + m], b->f[SP_U[s] + m], sizeof(double) * b->V);
        for (int a = 0; a < 3; ++a) memcpy(b->f[S_B0 + a], b->f[S_B + a], sizeof(double) * b->V);
    }
}

/* ------------------------------------------------------------------------- */
/* implicit resistive correction (Z4_Mercury/src/implicit/Mercury_ImplicitResistive.cpp) */
/* ------------------------------------------------------------------------- */

/* CTOperators::CurlEdgeToFace (Z4_Mercury/src/operators/CTOperators.cpp:6-57):
 * face = mult * circulation of the edge 1-form, owned faces */
static void curl_e2f(Orc *o, int se, int sf, double mult) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        int hi[3];
#define EE(a, i, j, k) F(b, se + (a), i, j, k)
        owned_hi(b, L_FX, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) F(b, sf, i, j, k) = mult * ((EE(1, i, j, k) - EE(1, i, j, k + 1)) + (EE(2, i, j + 1, k) - EE(2, i, j, k)));
        owned_hi(b, L_FY, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) F(b, sf + 1, i, j, k) = mult * ((EE(0, i, j, k + 1) - EE(0, i, j, k)) + (EE(2, i, j, k) - EE(2, i + 1, j, k)));
        owned_hi(b, L_FZ, hi);
        PAR2
        for (int i = 0; i < hi[0]; ++i)
            for (int j = 0; j < hi[1]; ++j)
                for (int k = 0; k < hi[2]; ++k) F(b, sf + 2, i, j, k) = mult * ((EE(0, i, j, k) - EE(0, i, j + 1, k)) + (EE(1, i + 1, j, k) - EE(1, i, j, k)));
#undef EE
    }
}

/* MercuryResistivityShape_ (Z4_Mercury/src/physics/MercurySolver_Tools.cpp:50-69);
 * profile 1: exp(-(r - r_inner) / width), an extension with no reference counterpart */
static double resist_shape(const OrcPhys *P, double r) {
    if (P->resist_profile == 1.0) return exp(-(r - P->resist_r_inner) / P->resist_width);
    const double shape = 0.5 * (tanh((r - P->resist_r_inner) / P->resist_width) - tanh((r - P->resist_r_outer) / P->resist_width));
    if (!isfinite(shape) || shape <= 1.0e-10) return 0.0;
    return dmin(1.0, shape);
}

/* Unknown edges and their eta (BuildImplicitResistiveEdgeDofMap_ :165-217,
 * ImplicitResistiveEtaAtEdge_ :219-252): owned edges that own their periodic
 * class, with dl > 1e-12 and eta = inver_Rem * shape(|edge midpoint|) > 0.
 * Two kinds are left out of the unknowns because the exact solution does not
 * depend on them: edges with alpha = 0 (their operator row is the identity and
 * their right-hand side eta*alpha*(...) is 0, so they solve to 0) and edges of
 * a far-field boundary layer (the Eres far-field handler overwrites them before
 * they reach any face).  Writes eta into S_ETAR (0 off the unknowns). */
static void resist_dofs(Orc *o) {
    const OrcPhys *P = &o->P;
    /* non-owner members of the edge classes */
    char **alias[3];
    for (int a = 0; a < 3; ++a) {
        alias[a] = (char **)calloc(o->nb, sizeof(char *));
        for (int ib = 0; ib < o->nb; ++ib) alias[a][ib] = (char *)calloc(o->b[ib].V, 1);
    }
    for (int L = L_EX; L <= L_EZ; ++L)
        for (int c = 0; c < o->ncls[L]; ++c)
            for (long t = o->cls_start[L][c] + 1; t < o->cls_start[L][c + 1]; ++t)
                alias[o->cls_ax[L][t]][o->cls_b[L][t]][o->cls_q[L][t]] = 1;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int a = 0; a < 3; ++a) {
            const int loc = L_EX + a;
            memset(b->f[S_ETAR + a], 0, sizeof(double) * b->V);
            int hi[3];
            owned_hi(b, loc, hi);
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) {
                        if (alias[a][ib][IX(b, i, j, k)]) continue;
                        if (fabs(F(b, S_DL + a, i, j, k)) <= 1.0e-12) continue;
                        const int ir = i + (a == 0), jr = j + (a == 1), kr = k + (a == 2);
                        const double xm = 0.5 * (XN(b, i, j, k) + XN(b, ir, jr, kr));
                        const double ym = 0.5 * (YN(b, i, j, k) + YN(b, ir, jr, kr));
                        const double zm = 0.5 * (ZN(b, i, j, k) + ZN(b, ir, jr, kr));
                        const double eta = P->inver_Rem * resist_shape(P, sqrt(xm * xm + ym * ym + zm * zm));
                        if (!(eta > 0.0) || F(b, S_ALPHA + a, i, j, k) == 0.0) continue;
                        F(b, S_ETAR + a, i, j, k) = eta;
                    }
            /* far-field boundary layers */
            for (int f = 0; f < 6; ++f) {
                if (b->bc[f] != BC_FARFIELD) continue;
                int dir = (f / 2 + 1) * ((f & 1) ? 1 : -1), lo[3], sh[3];
                inner_slab(b, loc, dir, lo, sh);
                for (int i = lo[0]; i < sh[0]; ++i)
                    for (int j = lo[1]; j < sh[1]; ++j)
                        for (int k = lo[2]; k < sh[2]; ++k) F(b, S_ETAR + a, i, j, k) = 0.0;
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        for (int ib = 0; ib < o->nb; ++ib) free(alias[a][ib]);
        free(alias[a]);
    }
}

typedef struct {
    int n;
    int *blk, *ax, *i, *j, *k;
} DofList;

static void dof_list(Orc *o, DofList *d) {
    d->n = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int n = 0;
        for (int ib = 0; ib < o->nb; ++ib) {
            OBlock *b = &o->b[ib];
            for (int a = 0; a < 3; ++a) {
                int hi[3];
                owned_hi(b, L_EX + a, hi);
                for (int i = 0; i < hi[0]; ++i)
                    for (int j = 0; j < hi[1]; ++j)
                        for (int k = 0; k < hi[2]; ++k) {
                            if (!(F(b, S_ETAR + a, i, j, k) > 0.0)) continue;
                            if (pass) d->blk[n] = ib, d->ax[n] = a, d->i[n] = i, d->j[n] = j, d->k[n] = k;
                            ++n;
                        }
            }
        }
        if (!pass) {
            d->n = n;
            d->blk = (int *)malloc(sizeof(int) * (n + 1));
            d->ax = (int *)malloc(sizeof(int) * (n + 1));
            d->i = (int *)malloc(sizeof(int) * (n + 1));
            d->j = (int *)malloc(sizeof(int) * (n + 1));
            d->k = (int *)malloc(sizeof(int) * (n + 1));
        }
    }
}
static void dof_free(DofList *d) { free(d->blk); free(d->ax); free(d->i); free(d->j); free(d->k); }
#define DOF(o, d, s, n) F(&(o)->b[(d)->blk[n]], (s) + (d)->ax[n], (d)->i[n], (d)->j[n], (d)->k[n])

/* MatMultImplicitResistive_ (Mercury_ImplicitResistive.cpp:406-465):
 * Eres <- x on the unknowns (0 elsewhere), Sync(Eres), dB = CurlEdgeToFace(Eres)
 * * (-dt), Sync(dB), dJ = alpha * CurlAdjFaceToEdge(beta * dB), alias mean,
 * Sync(dJ), y = Eres - eta * dJ on the unknowns.  lin = 1 applies the far-field
 * Eres handler with a zero field: the linear part of the (affine) operator. */
static void res_apply(Orc *o, const DofList *d, const double *x, double *y, double dt, int lin) {
    for (int ib = 0; ib < o->nb; ++ib)
        for (int a = 0; a < 3; ++a) {
            memset(o->b[ib].f[S_ERES + a], 0, sizeof(double) * o->b[ib].V);
            memset(o->b[ib].f[S_DB + a], 0, sizeof(double) * o->b[ib].V);
            memset(o->b[ib].f[S_DJ + a], 0, sizeof(double) * o->b[ib].V);
        }
    for (int n = 0; n < d->n; ++n) DOF(o, d, S_ERES, n) = x[n];
    sync_group(o, lin ? G_ERES0 : G_ERES);
    curl_e2f(o, S_ERES, S_DB, -dt);
    sync_group(o, G_DB);
    adj_curl(o, S_DB, S_DJ);
    alias_reduce_edges(o, S_DJ);
    singular_current(o, S_DB, S_DJ);
    sync_group(o, G_DJ);
    for (int n = 0; n < d->n; ++n) y[n] = DOF(o, d, S_ERES, n) - DOF(o, d, S_ETAR, n) * DOF(o, d, S_DJ, n);
}

/* diagonal of the operator for the Jacobi preconditioner: 1 + dt*eta*alpha*(sum
 * of beta over the four faces around the edge) */
static double res_diag(Orc *o, const DofList *d, int n, double dt) {
    OBlock *b = &o->b[d->blk[n]];
    const int a = d->ax[n], i = d->i[n], j = d->j[n], k = d->k[n];
    double s;
    if (a == 0) s = F(b, S_BETA + 1, i, j, k - 1) + F(b, S_BETA + 1, i, j, k) + F(b, S_BETA + 2, i, j, k) + F(b, S_BETA + 2, i, j - 1, k);
    else if (a == 1) s = F(b, S_BETA + 0, i, j, k) + F(b, S_BETA + 0, i, j, k - 1) + F(b, S_BETA + 2, i - 1, j, k) + F(b, S_BETA + 2, i, j, k);
    else s = F(b, S_BETA + 0, i, j - 1, k) + F(b, S_BETA + 0, i, j, k) + F(b, S_BETA + 1, i, j, k) + F(b, S_BETA + 1, i - 1, j, k);
    return 1.0 + dt * (DOF(o, d, S_ETAR, n) * F(b, S_ALPHA + a, i, j, k)) * s;
}

/* SolveImplicitResistiveDiffusion_ (Mercury_ImplicitResistive.cpp:467-547) and
 * ApplyImplicitResistiveUpdate_ (:549-575).  The system A(x) = b, b = eta*J(B*),
 * is solved by conjugate gradients on D^-1 A_lin x = D^-1 (b - A(0)), D =
 * diag(eta*alpha): D^-1 A_lin = D^-1 + dt C^T diag(beta) C is symmetric positive
 * definite.  Jacobi preconditioner; stop when |b - A(x)|_2 <= max(rtol |b|_2,
 * atol).  The previous solution is the initial guess.  The reference runs
 * PETSc GMRES without preconditioner on the same equation. */
static int res_solve(Orc *o, double dt, int *its_out, double *rel_out) {
    const OrcPhys *P = &o->P;
    DofList d;
    dof_list(o, &d);
    const int n = d.n;
    double *b = calloc(n + 1, sizeof(double)), *x = calloc(n + 1, sizeof(double));
    double *r = calloc(n + 1, sizeof(double)), *z = calloc(n + 1, sizeof(double)), *p = calloc(n + 1, sizeof(double));
    double *Ap = calloc(n + 1, sizeof(double)), *dg = calloc(n + 1, sizeof(double)), *w = calloc(n + 1, sizeof(double));
    /* Calc_J_Edge on B* (the synchronised state after the explicit update) */
    calc_jedge(o);
    alias_reduce_edges(o, S_J);
    singular_current(o, S_B, S_J);
    sync_group(o, G_JEDGE);
    for (int q = 0; q < n; ++q) {
        b[q] = 0.0 + DOF(o, &d, S_ETAR, q) * DOF(o, &d, S_J, q);
        x[q] = DOF(o, &d, S_XRES, q);
        dg[q] = res_diag(o, &d, q, dt);
        w[q] = 1.0 / (DOF(o, &d, S_ETAR, q) * F(&o->b[d.blk[q]], S_ALPHA + d.ax[q], d.i[q], d.j[q], d.k[q]));
    }
    res_apply(o, &d, x, Ap, dt, 0);
    double bn = 0.0, rn = 0.0, rz = 0.0;
    for (int q = 0; q < n; ++q) {
        r[q] = b[q] - Ap[q];
        z[q] = r[q] / dg[q];
        p[q] = z[q];
        bn += b[q] * b[q];
        rn += r[q] * r[q];
        rz += r[q] * z[q] * w[q];
    }
    bn = sqrt(bn);
    rn = sqrt(rn);
    const double tol = dmax(P->resist_rtol * bn, P->resist_atol);
    int its = 0, rc = 0;
    while (rn > tol) {
        if (its >= (int)P->resist_max_it) { rc = -3; break; }
        res_apply(o, &d, p, Ap, dt, 1);
        double pAp = 0.0;
        for (int q = 0; q < n; ++q) pAp += p[q] * Ap[q] * w[q];
        const double al = rz / pAp;
        double rzn = 0.0;
        rn = 0.0;
        for (int q = 0; q < n; ++q) {
            x[q] += al * p[q];
            r[q] -= al * Ap[q];
            z[q] = r[q] / dg[q];
            rn += r[q] * r[q];
            rzn += r[q] * z[q] * w[q];
        }
        rn = sqrt(rn);
        const double be = rzn / rz;
        rz = rzn;
        for (int q = 0; q < n; ++q) p[q] = z[q] + be * p[q];
        ++its;
    }
    if (its_out) *its_out = its;
    if (rel_out) *rel_out = bn > 0.0 ? rn / bn : rn;
    for (int ib = 0; ib < o->nb; ++ib)
        for (int a = 0; a < 3; ++a) memset(o->b[ib].f[S_XRES + a], 0, sizeof(double) * o->b[ib].V);
    for (int q = 0; q < n; ++q) DOF(o, &d, S_XRES, q) = x[q];
    /* UnpackVecToImplicitEres_ + ApplyImplicitResistiveUpdate_: Eres = x (synced),
     * RHS_Bres = -CurlEdgeToFace(Eres), B += dt * RHS_Bres on owned faces, Sync(Bface) */
    for (int ib = 0; ib < o->nb; ++ib)
        for (int a = 0; a < 3; ++a) memset(o->b[ib].f[S_ERES + a], 0, sizeof(double) * o->b[ib].V);
    for (int q = 0; q < n; ++q) DOF(o, &d, S_ERES, q) = x[q];
    sync_group(o, G_ERES);
    curl_e2f(o, S_ERES, S_RHSBR, -1.0);
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *bk = &o->b[ib];
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(bk, L_FX + a, hi);
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) F(bk, S_B + a, i, j, k) += dt * F(bk, S_RHSBR + a, i, j, k);
        }
    }
    sync_group(o, G_BFACE);
    free(b); free(x); free(r); free(z); free(p); free(Ap); free(dg); free(w);
    dof_free(&d);
    return rc;
}

/* the operator on fields, for parity tests: x from the generic edge 1-form
 * slots on the unknowns, y written back there (0 elsewhere) */
static void res_apply_fields(Orc *o, double dt, int lin) {
    DofList d;
    dof_list(o, &d);
    double *x = calloc(d.n + 1, sizeof(double)), *y = calloc(d.n + 1, sizeof(double));
    for (int q = 0; q < d.n; ++q) x[q] = DOF(o, &d, S_EDGEF, q);
    res_apply(o, &d, x, y, dt, lin);
    for (int ib = 0; ib < o->nb; ++ib)
        for (int a = 0; a < 3; ++a) memset(o->b[ib].f[S_EDGEF + a], 0, sizeof(double) * o->b[ib].V);
    for (int q = 0; q < d.n; ++q) DOF(o, &d, S_EDGEF, q) = y[q];
    free(x); free(y);
    dof_free(&d);
}

/* one evaluation of the stage operator and Euler update:
 * StepOnce (MercurySolver_Driver.cpp:21-61) without the dt computation */
static int stage(Orc *o, double dt, int resist) {
    scheme_u(o);
    add_sources(o);
    induction_ct(o);
    apply_update(o, dt);
    sync_group(o, G_UCELL);
    sync_group(o, G_BFACE);
    int rc = 0;
    if (resist) rc = res_solve(o, dt, NULL, NULL);
    update_derived(o);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* C ABI for the tests                                                        */
/* ------------------------------------------------------------------------- */

/* Blocks with 1:1 interface connections: for every face with code
 * BC_INTERFACE, nbr[6*ib+f] names the neighbour block and trans[9*(6*ib+f)..]
 * its IndexTransform (perm[3], sign[3], offset[3]: neighbour node
 * index[perm[a]] = sign[a] * node index[a] + offset[a],
 * include/2_topology/TopologyTypes.h:93-98).  Supported: translations
 * (perm = identity, sign = +1) that glue the face to the opposite face of the
 * neighbour over its whole extent, declared from both sides, with at least
 * as many cells as ghost layers along the connection.  Returns NULL on
 * anything else. */
/* ------------------------------------------------------------------------- */
/* patch connections: shared-DOF classes and the context                      */
/* ------------------------------------------------------------------------- */
/* Equivalence classes through patches (TOPO::detail::build_equivalence,
 * src/2_topology/Topology.cpp:280-323): nodes joined by every patch's node map
 * form classes whose representative is the smallest EntityKey (block, i, j, k)
 * (TopologyNode.cpp:180-235).  An edge's key is its two endpoint
 * representatives in ascending order, its sign +1 when they come in that order
 * (TOPO::make_edge_key, TopologyQuery.cpp:104-126); a face's key is its four
 * sorted corner representatives (make_face_key, :128-184) and its sign
 * compares its boundary-edge stencil with the owner's: +1 equal, -1 opposite
 * (orient_face_members_to_owner, TopologyFaceEquiv.cpp:298-351; boundary of a
 * face, include/2_topology/LocalIncidence.h:44-79).  Owner = smallest member
 * (block, i, j, k, axis); edge owners keep their own sign, face owners +1
 * (TopologyEdgeEquiv.cpp:176-262, TopologyFaceEquiv.cpp:353-461).  Only
 * classes of two or more members are kept: the others take part in no copy. */
typedef struct { long k0, k1, k2, k3; int b, i, j, k, ax, sg; long st[4][3]; } PEnt; /* st: (edge key lo, hi, coefficient) */
static int pent_cmp(const void *x, const void *y) {
    const PEnt *a = (const PEnt *)x, *b = (const PEnt *)y;
    const long ka[8] = {a->k0, a->k1, a->k2, a->k3, a->b, a->i, a->j, a->k};
    const long kb[8] = {b->k0, b->k1, b->k2, b->k3, b->b, b->i, b->j, b->k};
    for (int t = 0; t < 8; ++t)
        if (ka[t] != kb[t]) return ka[t] < kb[t] ? -1 : 1;
    return (a->ax > b->ax) - (a->ax < b->ax);
}
static int st_cmp(const void *x, const void *y) {
    const long *a = (const long *)x, *b = (const long *)y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    return (a[1] > b[1]) - (a[1] < b[1]);
}
static int build_classes_patched(Orc *o) {
    for (int L = 0; L < 8; ++L) {
        o->ncls[L] = 0;
        o->cls_start[L] = (long *)calloc(1, sizeof(long));
        o->cls_b[L] = NULL;
        o->cls_q[L] = NULL;
        o->cls_ax[L] = NULL;
        o->cls_sg[L] = NULL;
    }
    long *nbase = (long *)calloc(o->nb + 1, sizeof(long));
    for (int ib = 0; ib < o->nb; ++ib) nbase[ib + 1] = nbase[ib] + (long)(o->b[ib].m[0] + 1) * (o->b[ib].m[1] + 1) * (o->b[ib].m[2] + 1);
#define NODE(ib, i, j, k) (nbase[ib] + (((long)(i) * (o->b[ib].m[1] + 1) + (j)) * (o->b[ib].m[2] + 1) + (k)))
    const long nn = nbase[o->nb];
    long *par = (long *)malloc(sizeof(long) * nn);
    long *cnt = (long *)calloc(nn, sizeof(long));
    for (long x = 0; x < nn; ++x) par[x] = x;
    for (int p = 0; p < o->npatch; ++p) {
        const int *P = o->patch + PT_N * p;
        int q[3], s[3];
        for (q[0] = P[PT_LO]; q[0] < P[PT_HI]; ++q[0])
            for (q[1] = P[PT_LO + 1]; q[1] < P[PT_HI + 1]; ++q[1])
                for (q[2] = P[PT_LO + 2]; q[2] < P[PT_HI + 2]; ++q[2]) {
                    tr_map(P + PT_PERM, P + PT_SIGN, P + PT_OFF, q, s);
                    const long rx = uf_find(par, NODE(P[PT_BLOCK], q[0], q[1], q[2])), ry = uf_find(par, NODE(P[PT_NB], s[0], s[1], s[2]));
                    if (rx != ry) par[rx < ry ? ry : rx] = rx < ry ? rx : ry;
                }
    }
    for (long x = 0; x < nn; ++x) cnt[uf_find(par, x)] += 1;
    for (int dim = 0; dim < 2; ++dim) { /* 0: edges, 1: faces */
        PEnt *e = NULL;
        long ne = 0, cap = 0;
        for (int ib = 0; ib < o->nb; ++ib) {
            const OBlock *b = &o->b[ib];
            for (int ax = 0; ax < 3; ++ax) {
                const int L = (dim ? L_FX : L_EX) + ax;
                int hi[3];
                owned_hi(b, L, hi);
                for (int i = 0; i < hi[0]; ++i)
                    for (int j = 0; j < hi[1]; ++j)
                        for (int k = 0; k < hi[2]; ++k) {
                            PEnt x;
                            memset(&x, 0, sizeof x);
                            x.b = ib, x.i = i, x.j = j, x.k = k, x.ax = ax, x.sg = 1;
                            if (!dim) {
                                const long r0 = uf_find(par, NODE(ib, i, j, k));
                                if (cnt[r0] <= 1) continue;
                                const long r1 = uf_find(par, NODE(ib, i + (ax == 0), j + (ax == 1), k + (ax == 2)));
                                if (cnt[r1] <= 1 || r0 == r1) continue;
                                x.k0 = r0 < r1 ? r0 : r1;
                                x.k1 = r0 < r1 ? r1 : r0;
                                x.sg = r0 < r1 ? 1 : -1;
                            } else {
                                /* corners n0..n3 and the four boundary edges (start, axis, incidence) */
                                const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
                                long c[4];
                                int ok = 1;
                                for (int n = 0; n < 4 && ok; ++n) {
                                    int q[3] = {i, j, k};
                                    q[t1] += n & 1;
                                    q[t2] += n >> 1;
                                    c[n] = uf_find(par, NODE(ib, q[0], q[1], q[2]));
                                    ok = cnt[c[n]] > 1;
                                }
                                if (!ok) continue;
                                long srt[4] = {c[0], c[1], c[2], c[3]};
                                for (int u = 1; u < 4; ++u)
                                    for (int w = u; w > 0 && srt[w] < srt[w - 1]; --w) {
                                        const long z = srt[w];
                                        srt[w] = srt[w - 1];
                                        srt[w - 1] = z;
                                    }
                                x.k0 = srt[0], x.k1 = srt[1], x.k2 = srt[2], x.k3 = srt[3];
                                /* boundary_of_face: (+ t1-edge at start, + t2-edge at start+e_t1, - t1-edge at start+e_t2, - t2-edge at start) for
                                 * Xi (t1 eta, t2 zeta) and Zeta (t1 xi, t2 eta); Eta lists zeta first: (+ zeta at start, + xi at +e_zeta,
                                 * - zeta at +e_xi, - xi at start) */
                                int ea[4], es[4][3], ei[4];
                                const int u1 = ax == 1 ? 2 : t1, u2 = ax == 1 ? 0 : t2;
                                for (int n = 0; n < 4; ++n)
                                    for (int a = 0; a < 3; ++a) es[n][a] = (a == 0 ? i : (a == 1 ? j : k));
                                ea[0] = u1, ei[0] = 1;
                                ea[1] = u2, ei[1] = 1, es[1][u1] += 1;
                                ea[2] = u1, ei[2] = -1, es[2][u2] += 1;
                                ea[3] = u2, ei[3] = -1;
                                long st[4][3];
                                int ns = 0;
                                for (int n = 0; n < 4; ++n) {
                                    const long a0 = uf_find(par, NODE(ib, es[n][0], es[n][1], es[n][2]));
                                    const long a1 = uf_find(par, NODE(ib, es[n][0] + (ea[n] == 0), es[n][1] + (ea[n] == 1), es[n][2] + (ea[n] == 2)));
                                    const long lo = a0 < a1 ? a0 : a1, hi2 = a0 < a1 ? a1 : a0;
                                    const long coef = ei[n] * (a0 < a1 ? 1 : -1);
                                    int hit = -1;
                                    for (int w = 0; w < ns; ++w)
                                        if (st[w][0] == lo && st[w][1] == hi2) hit = w;
                                    if (hit < 0) {
                                        st[ns][0] = lo, st[ns][1] = hi2, st[ns][2] = coef;
                                        ++ns;
                                    } else {
                                        st[hit][2] += coef;
                                    }
                                }
                                int w2 = 0;
                                for (int w = 0; w < ns; ++w)
                                    if (st[w][2] != 0) {
                                        for (int z = 0; z < 3; ++z) st[w2][z] = st[w][z];
                                        ++w2;
                                    }
                                qsort(st, w2, sizeof st[0], st_cmp);
                                for (int w = 0; w < 4; ++w)
                                    for (int z = 0; z < 3; ++z) x.st[w][z] = w < w2 ? st[w][z] : 0;
                            }
                            if (ne == cap) {
                                cap = cap ? 2 * cap : 1024;
                                e = (PEnt *)realloc(e, sizeof(PEnt) * cap);
                            }
                            e[ne++] = x;
                        }
            }
        }
        qsort(e, ne, sizeof(PEnt), pent_cmp);
        const int L = dim ? L_FX : L_EX;
        long *st = (long *)malloc(sizeof(long) * (ne + 1));
        int *cb = (int *)malloc(sizeof(int) * (ne + 1));
        long *cq = (long *)malloc(sizeof(long) * (ne + 1));
        int *ca = (int *)malloc(sizeof(int) * (ne + 1));
        signed char *cs = (signed char *)malloc(ne + 1);
        int nc = 0;
        long nm = 0;
        for (long u = 0; u < ne;) {
            long v = u + 1;
            while (v < ne && e[v].k0 == e[u].k0 && e[v].k1 == e[u].k1 && e[v].k2 == e[u].k2 && e[v].k3 == e[u].k3) ++v;
            if (v - u >= 2) {
                st[nc++] = nm;
                for (long w = u; w < v; ++w) {
                    int sg = e[w].sg;
                    if (dim) {
                        int same = 1, opp = 1;
                        for (int z = 0; z < 4; ++z) {
                            same &= e[w].st[z][0] == e[u].st[z][0] && e[w].st[z][1] == e[u].st[z][1] && e[w].st[z][2] == e[u].st[z][2];
                            opp &= e[w].st[z][0] == e[u].st[z][0] && e[w].st[z][1] == e[u].st[z][1] && e[w].st[z][2] == -e[u].st[z][2];
                        }
                        if (!same && !opp) {
                            free(e), free(st), free(cb), free(cq), free(ca), free(cs), free(par), free(cnt), free(nbase);
                            return 0;
                        }
                        sg = same ? 1 : -1;
                    }
                    cb[nm] = e[w].b;
                    cq[nm] = IX(&o->b[e[w].b], e[w].i, e[w].j, e[w].k);
                    ca[nm] = e[w].ax;
                    cs[nm] = (signed char)sg;
                    ++nm;
                }
            }
            u = v;
        }
        st[nc] = nm;
        free(o->cls_start[L]);
        o->ncls[L] = nc;
        o->cls_start[L] = st;
        o->cls_b[L] = cb;
        o->cls_q[L] = cq;
        o->cls_ax[L] = ca;
        o->cls_sg[L] = cs;
        free(e);
    }
#undef NODE
    free(par);
    free(cnt);
    free(nbase);
    return 1;
}

void *orc_create_conn(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans, const double *nodes,
                      const OrcPhys *P);
void *orc_create(int nb, const int *dims, int g, const int *bc, const double *nodes, const OrcPhys *P) {
    return orc_create_conn(nb, dims, g, bc, NULL, NULL, nodes, P);
}
static int conn_valid(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans) {
    if (nb <= 0) return 0;
    for (int ib = 0; ib < nb; ++ib)
        for (int f = 0; f < 6; ++f) {
            const int c = bc[6 * ib + f];
            if (c < 0 || c > 3) return 0;
            if (c == BC_PERIODIC && bc[6 * ib + (f ^ 1)] != BC_PERIODIC) return 0;
            if (c != BC_INTERFACE) continue;
            if (!nbr || !trans) return 0;
            const int n = nbr[6 * ib + f], a = f / 2, fo = f ^ 1;
            if (n < 0 || n >= nb || bc[6 * n + fo] != BC_INTERFACE || nbr[6 * n + fo] != ib) return 0;
            const int *t = trans + 9 * (6 * ib + f), *u = trans + 9 * (6 * n + fo);
            for (int d = 0; d < 3; ++d) {
                if (t[d] != d || t[3 + d] != 1 || t[6 + d] != -u[6 + d]) return 0;
                if (d != a && (t[6 + d] != 0 || dims[3 * ib + d] != dims[3 * n + d])) return 0;
            }
            if (t[6 + a] != ((f & 1) ? -dims[3 * ib + a] : dims[3 * n + a])) return 0;
            if (dims[3 * ib + a] < g || dims[3 * n + a] < g) return 0;
        }
    return 1;
}
void *orc_create_conn(int nb, const int *dims, int g, const int *bc, const int *nbr, const int *trans, const double *nodes,
                      const OrcPhys *P) {
    if (!conn_valid(nb, dims, g, bc, nbr, trans)) return NULL;
    Orc *o = (Orc *)calloc(1, sizeof(Orc));
    o->nb = nb;
    o->b = (OBlock *)calloc(nb, sizeof(OBlock));
    o->P = *P;
    long noff = 0;
    for (int ib = 0; ib < nb; ++ib) {
        OBlock *b = &o->b[ib];
        b->g = g;
        for (int a = 0; a < 3; ++a) {
            b->m[a] = dims[3 * ib + a];
            b->N[a] = b->m[a] + 1 + 2 * g;
            b->Nn[a] = b->m[a] + 1 + 2 * (g + 1);
        }
        for (int f = 0; f < 6; ++f) {
            b->bc[f] = bc[6 * ib + f];
            b->nbr[f] = -1;
            if (b->bc[f] == BC_INTERFACE) {
                b->nbr[f] = nbr[6 * ib + f];
                for (int c = 0; c < 3; ++c) b->off[f][c] = trans[9 * (6 * ib + f) + 6 + c];
            }
        }
        b->V = (long)b->N[0] * b->N[1] * b->N[2];
        b->Vn = (long)b->Nn[0] * b->Nn[1] * b->Nn[2];
        b->x = (double *)calloc(b->Vn, sizeof(double));
        b->y = (double *)calloc(b->Vn, sizeof(double));
        b->z = (double *)calloc(b->Vn, sizeof(double));
        for (int s = 0; s < S_NSLOT; ++s) b->f[s] = (double *)calloc(b->V, sizeof(double));
        /* owned nodes (m+1)^3, k fastest, x then y then z planes */
        long nn = (long)(b->m[0] + 1) * (b->m[1] + 1) * (b->m[2] + 1);
        long q = 0;
        for (int i = 0; i <= b->m[0]; ++i)
            for (int j = 0; j <= b->m[1]; ++j)
                for (int k = 0; k <= b->m[2]; ++k, ++q) {
                    XN(b, i, j, k) = nodes[noff + q];
                    YN(b, i, j, k) = nodes[noff + nn + q];
                    ZN(b, i, j, k) = nodes[noff + 2 * nn + q];
                }
        noff += 3 * nn;
        add_ghostmesh(b);
        dual_grid(b);
        metric(b);
        recon_weights(b);
        edge_statics(b, &o->P);
    }
    build_classes(o);
    resist_dofs(o);
    return o;
}

void orc_destroy(void *h) {
    Orc *o = (Orc *)h;
    if (!o) return;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        free(b->x); free(b->y); free(b->z);
        for (int s = 0; s < S_NSLOT; ++s) free(b->f[s]);
    }
    for (int L = 0; L < 8; ++L) {
        free(o->cls_start[L]);
        free(o->cls_b[L]);
        free(o->cls_q[L]);
        free(o->cls_ax[L]);
        free(o->cls_sg[L]);
    }
    free(o->patch);
    for (int l = 0; l < 4; ++l) free(o->reg[l]);
    if (o->singp) {
        OSingReg *S = o->singp;
        free(S->e), free(S->cb), free(S->cax), free(S->cor), free(S->cq), free(S->cw);
        free(S->fb), free(S->fax), free(S->fq), free(S->fcoef), free(S);
    }
    free(o->b);
    free(o);
}


/* A context whose block faces connect through patches (18 ints each, see
 * PT_*): every face with code BC_INTERFACE must be covered by the node boxes of
 * its patches; periodicity is a patch from a block to itself.  rule 0 maps
 * staggered DOFs as the reference does, rule 1 by position (region_src).
 * NULL when the patches are not valid. */
static int patches_valid(int nb, const int *dims, int g, const int *bc, int npatch, const int *pt) {
    (void)g;
    if (nb <= 0 || nb > 64 || npatch < 0 || (npatch > 0 && !pt)) return 0;
    for (int ib = 0; ib < nb; ++ib)
        for (int f = 0; f < 6; ++f)
            if (bc[6 * ib + f] < BC_FARFIELD || bc[6 * ib + f] > BC_INTERFACE) return 0;
    for (int p = 0; p < npatch; ++p) {
        const int *P = pt + PT_N * p;
        const int b = P[PT_BLOCK], d = P[PT_DIR], n = P[PT_NB];
        if (b < 0 || b >= nb || n < 0 || n >= nb || d == 0 || abs(d) > 3) return 0;
        const int *m = dims + 3 * b, *mn = dims + 3 * n;
        int seen[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) {
            const int q = P[PT_PERM + a];
            if (q < 0 || q > 2 || seen[q]++) return 0;
            if (P[PT_SIGN + a] != 1 && P[PT_SIGN + a] != -1) return 0;
            if (P[PT_LO + a] < 0 || P[PT_HI + a] <= P[PT_LO + a] || P[PT_HI + a] > m[a] + 1) return 0;
        }
        if (face_dir_of(P + PT_LO, P + PT_HI, m) != d) return 0;
        if (bc[6 * b + 2 * (abs(d) - 1) + (d > 0)] != BC_INTERFACE) return 0;
        int lo[3], hi[3];
        tr_box(P + PT_PERM, P + PT_SIGN, P + PT_OFF, P + PT_LO, P + PT_HI, lo, hi);
        for (int a = 0; a < 3; ++a)
            if (lo[a] < 0 || hi[a] > mn[a] + 1) return 0;
        const int dn = face_dir_of(lo, hi, mn);
        if (dn == 0 || bc[6 * n + 2 * (abs(dn) - 1) + (dn > 0)] != BC_INTERFACE) return 0;
    }
    /* every node of an interface face lies in a patch of that face */
    for (int ib = 0; ib < nb; ++ib)
        for (int f = 0; f < 6; ++f) {
            if (bc[6 * ib + f] != BC_INTERFACE) continue;
            const int a = f / 2, d = (f & 1) ? a + 1 : -(a + 1), *m = dims + 3 * ib;
            const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
            for (int u = 0; u <= m[t1]; ++u)
                for (int w = 0; w <= m[t2]; ++w) {
                    int q[3];
                    q[a] = (f & 1) ? m[a] : 0;
                    q[t1] = u;
                    q[t2] = w;
                    int cov = 0;
                    for (int p = 0; p < npatch && !cov; ++p) {
                        const int *P = pt + PT_N * p;
                        if (P[PT_BLOCK] != ib || P[PT_DIR] != d) continue;
                        cov = 1;
                        for (int c = 0; c < 3; ++c) cov &= P[PT_LO + c] <= q[c] && q[c] < P[PT_HI + c];
                    }
                    if (!cov) return 0;
                }
        }
    return 1;
}
void *orc_create_patched(int nb, const int *dims, int g, const int *bc, int npatch, const int *pt, int rule, const double *nodes,
                         const OrcPhys *P) {
    if (!patches_valid(nb, dims, g, bc, npatch, pt) || (rule != 0 && rule != 1)) return NULL;
    /* faces carry no per-face connection: the context is built without one and
     * the patch halo and classes replace the translation ones */
    int *bc0 = (int *)malloc(sizeof(int) * 6 * nb);
    for (int t = 0; t < 6 * nb; ++t) bc0[t] = bc[t] == BC_INTERFACE ? BC_FARFIELD : bc[t];
    Orc *o = (Orc *)orc_create_conn(nb, dims, g, bc0, NULL, NULL, nodes, P);
    free(bc0);
    if (!o) return NULL;
    for (int ib = 0; ib < nb; ++ib)
        for (int f = 0; f < 6; ++f) o->b[ib].bc[f] = bc[6 * ib + f];
    o->npatch = npatch;
    o->rule = rule;
    o->patch = (int *)malloc(sizeof(int) * PT_N * (npatch ? npatch : 1));
    memcpy(o->patch, pt, sizeof(int) * PT_N * npatch);
    build_patch_regions(o);
    for (int L = 0; L < 8; ++L) {
        free(o->cls_start[L]);
        free(o->cls_b[L]);
        free(o->cls_q[L]);
        free(o->cls_ax[L]);
        free(o->cls_sg[L]);
    }
    if (!build_classes_patched(o)) {
        orc_destroy(o);
        return NULL;
    }
    if (!build_singular(o)) {
        orc_destroy(o);
        return NULL;
    }
    resist_dofs(o);
    return o;
}

int orc_nslot(void) { return S_NSLOT; }

/* the singular-edge registry of a patch context: the entry count, and per
 * entry (owner block, i, j, k, axis, owner sign, cells, faces) and
 * (inverse Hodge, mean edge vector x, y, z) */
int orc_singular_count(void *h) {
    const Orc *o = (const Orc *)h;
    return o->singp ? o->singp->nsing : 0;
}
int orc_singular_entry(void *h, int n, int *iv, double *dv) {
    const Orc *o = (const Orc *)h;
    if (!o->singp || n < 0 || n >= o->singp->nsing) return -1;
    const OSing *E = &o->singp->e[n];
    int p[3];
    ijk_of(&o->b[E->ob], E->oq, p);
    iv[0] = E->ob, iv[1] = p[0], iv[2] = p[1], iv[3] = p[2], iv[4] = E->oax, iv[5] = E->osign;
    iv[6] = E->c1 - E->c0, iv[7] = E->f1 - E->f0;
    dv[0] = E->ih, dv[1] = E->cv[0], dv[2] = E->cv[1], dv[3] = E->cv[2];
    return 0;
}
long orc_block_volume(void *h, int ib) { return ((Orc *)h)->b[ib].V; }

/* slot lookup by reference field name (Z4_Mercury/include/0_MercuryFieldCatalog.h:70-202,
 * metric names of src/7_metric/Metric.cpp:12-76) */
int orc_slot(const char *name) {
    static const struct { const char *n; int s; } T[] = {
        {"U_H", S_UH}, {"U_Na", S_UN}, {"PV_H", S_PVH}, {"PV_Na", S_PVN}, {"U_plus", S_UP}, {"B_cell", S_BC},
        {"U_He", S_UHE}, {"PV_He", S_PVHE}, {"RHS_He", S_RHE},
        {"Bind_cell", S_BI}, {"J_cell", S_JC}, {"RHS_H", S_RH}, {"RHS_Na", S_RN}, {"Ehc", S_EHC}, {"hall_beta", S_HB},
        {"Jac", S_JAC}, {"Bcell_from_Bface_w", S_WB}, {"Jcell_from_Jedge_w", S_WJ}, {"Photo_rate", S_PHOTO},
        {"cell_center", S_CC}, {"divB", S_DIVB},
        {"B_xi", S_B}, {"B_eta", S_B + 1}, {"B_zeta", S_B + 2}, {"Badd_xi", S_BADD}, {"Badd_eta", S_BADD + 1}, {"Badd_zeta", S_BADD + 2},
        {"RHS_B_xi", S_RHSB}, {"RHS_B_eta", S_RHSB + 1}, {"RHS_B_zeta", S_RHSB + 2},
        {"JDxi", S_JD}, {"JDet", S_JD + 3}, {"JDze", S_JD + 6}, {"Area_xi", S_AREA}, {"Area_eta", S_AREA + 1}, {"Area_zeta", S_AREA + 2},
        {"beta_xi", S_BETA}, {"beta_eta", S_BETA + 1}, {"beta_zeta", S_BETA + 2},
        {"Eface_xi", S_EF}, {"Eface_eta", S_EF + 3}, {"Eface_zeta", S_EF + 6},
        {"E_xi", S_E}, {"E_eta", S_E + 1}, {"E_zeta", S_E + 2}, {"Ehall_xi", S_EH}, {"Ehall_eta", S_EH + 1}, {"Ehall_zeta", S_EH + 2},
        {"J_xi", S_J}, {"J_eta", S_J + 1}, {"J_zeta", S_J + 2}, {"dr_xi", S_DR}, {"dr_eta", S_DR + 3}, {"dr_zeta", S_DR + 6},
        {"dl_xi", S_DL}, {"dl_eta", S_DL + 1}, {"dl_zeta", S_DL + 2}, {"alpha_xi", S_ALPHA}, {"alpha_eta", S_ALPHA + 1}, {"alpha_zeta", S_ALPHA + 2},
        {"taper_xi", S_TAPER}, {"taper_eta", S_TAPER + 1}, {"taper_zeta", S_TAPER + 2},
        {"eta_local_xi", S_LETA}, {"eta_local_eta", S_LETA + 1}, {"eta_local_zeta", S_LETA + 2},
        {"node_scalar", S_NODEF}, {"edge_xi", S_EDGEF}, {"edge_eta", S_EDGEF + 1}, {"edge_zeta", S_EDGEF + 2},
        {"face_xi", S_FACEF}, {"face_eta", S_FACEF + 1}, {"face_zeta", S_FACEF + 2}, {"cell_scalar", S_CELLS},
        {"Eres_xi", S_ERES}, {"Eres_eta", S_ERES + 1}, {"Eres_zeta", S_ERES + 2}, {"dJ_xi", S_DJ}, {"dJ_eta", S_DJ + 1}, {"dJ_zeta", S_DJ + 2},
        {"eta_res_xi", S_ETAR}, {"eta_res_eta", S_ETAR + 1}, {"eta_res_zeta", S_ETAR + 2},
        {"Eres_guess_xi", S_XRES}, {"Eres_guess_eta", S_XRES + 1}, {"Eres_guess_zeta", S_XRES + 2},
        {"dB_xi", S_DB}, {"dB_eta", S_DB + 1}, {"dB_zeta", S_DB + 2},
        {"RHS_Bres_xi", S_RHSBR}, {"RHS_Bres_eta", S_RHSBR + 1}, {"RHS_Bres_zeta", S_RHSBR + 2}};
    for (unsigned q = 0; q < sizeof(T) / sizeof(T[0]); ++q)
        if (!strcmp(T[q].n, name)) return T[q].s;
    return -1;
}

int orc_get(void *h, int ib, int slot, double *out) {
    Orc *o = (Orc *)h;
    if (ib < 0 || ib >= o->nb || slot < 0 || slot >= S_NSLOT) return -1;
    memcpy(out, o->b[ib].f[slot], sizeof(double) * o->b[ib].V);
    return 0;
}
int orc_set(void *h, int ib, int slot, const double *in) {
    Orc *o = (Orc *)h;
    if (ib < 0 || ib >= o->nb || slot < 0 || slot >= S_NSLOT) return -1;
    memcpy(o->b[ib].f[slot], in, sizeof(double) * o->b[ib].V);
    return 0;
}
int orc_get_nodes(void *h, int ib, double *out) {
    Orc *o = (Orc *)h;
    OBlock *b = &o->b[ib];
    memcpy(out, b->x, sizeof(double) * b->Vn);
    memcpy(out + b->Vn, b->y, sizeof(double) * b->Vn);
    memcpy(out + 2 * b->Vn, b->z, sizeof(double) * b->Vn);
    return 0;
}

/* B = Bu . JD + d1(A) over the allocated face ranges; Badd = B_imf . JD.
 * (initial condition of the configurations; the face flux of a uniform field is
 * its dot product with the face area vector, as in DipoleField::Build_Badd_FaceFlux,
 * Z4_Mercury/include/operators/Dipole.h:89-374, for the IMF part) */
int orc_init_bface(void *h, const double *Bu, const double *const *A) {
    Orc *o = (Orc *)h;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        const double *Ax = A ? A[3 * ib] : NULL, *Ay = A ? A[3 * ib + 1] : NULL, *Az = A ? A[3 * ib + 2] : NULL;
#define AA(p, i, j, k) ((p) ? (p)[IX(b, i, j, k)] : 0.0)
        for (int a = 0; a < 3; ++a) {
            int lo[3], hi[3];
            alloc_box(b, L_FX + a, b->g, lo, hi);
            for (int i = lo[0]; i < hi[0]; ++i)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int k = lo[2]; k < hi[2]; ++k) {
                        double jd0 = F(b, S_JD + 3 * a, i, j, k), jd1 = F(b, S_JD + 3 * a + 1, i, j, k), jd2 = F(b, S_JD + 3 * a + 2, i, j, k);
                        double c;
                        if (a == 0) c = (AA(Ay, i, j, k) - AA(Ay, i, j, k + 1)) + (AA(Az, i, j + 1, k) - AA(Az, i, j, k));
                        else if (a == 1) c = (AA(Ax, i, j, k + 1) - AA(Ax, i, j, k)) + (AA(Az, i, j, k) - AA(Az, i + 1, j, k));
                        else c = (AA(Ax, i, j, k) - AA(Ax, i, j + 1, k)) + (AA(Ay, i + 1, j, k) - AA(Ay, i, j, k));
                        F(b, S_B + a, i, j, k) = (Bu[0] * jd0 + Bu[1] * jd1 + Bu[2] * jd2) + c;
                        F(b, S_BADD + a, i, j, k) = o->P.B_imf[0] * jd0 + o->P.B_imf[1] * jd1 + o->P.B_imf[2] * jd2;
                    }
        }
#undef AA
    }
    return 0;
}

/* DipoleField::A_of_dipole (Z4_Mercury/include/operators/Dipole.h:25-39): the
 * vector potential sum_d (m_d x r) / (|r|^2 + a_d^2)^(3/2), r = p - r0_d; dip
 * holds 7 nondimensional values per dipole: x0 y0 z0 mx my mz a.  Pinned by
 * the closed-form curl (3 r (r.m) - m (r^2 - 2 a^2)) / (r^2 + a^2)^(5/2) of the
 * face fluxes and by Mercury's equatorial surface field (tests). */
static void dipole_A(int nd, const double *dip, const double p[3], double A[3]) {
    A[0] = A[1] = A[2] = 0.0;
    for (int d = 0; d < nd; ++d) {
        const double *D = dip + 7 * d;
        const double r[3] = {p[0] - D[0], p[1] - D[1], p[2] - D[2]};
        double f = (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + D[6] * D[6];
        f = sqrt(f) * f;
        A[0] += (D[4] * r[2] - D[5] * r[1]) / f;
        A[1] += (D[5] * r[0] - D[3] * r[2]) / f;
        A[2] += (D[3] * r[1] - D[4] * r[0]) / f;
    }
}

/* DipoleField::Build_Badd_FaceFlux steps 1-2 (Dipole.h:128-324): the edge
 * 1-form A.dr (midpoint value of A times the edge vector) on the edge boxes of
 * ghost depth 1, then its curl on the face boxes of ghost depth 1, added to the
 * IMF flux already in Badd (step 3, :326-372; one rounded addition, so the
 * order of the two parts does not matter).  Faces beyond depth 1 keep the IMF
 * flux alone until Sync(Badd) fills them. */
int orc_add_dipole_badd(void *h, int nd, const double *dip) {
    Orc *o = (Orc *)h;
    if (nd <= 0) return 0;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        double *Ae[3];
        for (int a = 0; a < 3; ++a) {
            Ae[a] = (double *)calloc((size_t)b->V, sizeof(double));
            int lo[3], hi[3];
            alloc_box(b, L_EX + a, 1, lo, hi);
            for (int i = lo[0]; i < hi[0]; ++i)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int k = lo[2]; k < hi[2]; ++k) {
                        double x1[3], x2[3], xm[3], dr[3], A[3];
                        node(b, i, j, k, x1);
                        node(b, i + (a == 0), j + (a == 1), k + (a == 2), x2);
                        for (int c = 0; c < 3; ++c) {
                            xm[c] = (x1[c] + x2[c]) * 0.5;
                            dr[c] = x2[c] - x1[c];
                        }
                        dipole_A(nd, dip, xm, A);
                        Ae[a][IX(b, i, j, k)] = A[0] * dr[0] + A[1] * dr[1] + A[2] * dr[2];
                    }
        }
#define AE(c, i, j, k) Ae[c][IX(b, i, j, k)]
        for (int a = 0; a < 3; ++a) {
            int lo[3], hi[3];
            alloc_box(b, L_FX + a, 1, lo, hi);
            for (int i = lo[0]; i < hi[0]; ++i)
                for (int j = lo[1]; j < hi[1]; ++j)
                    for (int k = lo[2]; k < hi[2]; ++k) {
                        double c;
                        if (a == 0) c = (AE(1, i, j, k) - AE(1, i, j, k + 1)) + (AE(2, i, j + 1, k) - AE(2, i, j, k));
                        else if (a == 1) c = (AE(0, i, j, k + 1) - AE(0, i, j, k)) + (AE(2, i, j, k) - AE(2, i + 1, j, k));
                        else c = (AE(0, i, j, k) - AE(0, i, j + 1, k)) + (AE(1, i + 1, j, k) - AE(1, i, j, k));
                        F(b, S_BADD + a, i, j, k) = c + F(b, S_BADD + a, i, j, k);
                    }
        }
#undef AE
        for (int a = 0; a < 3; ++a) free(Ae[a]);
    }
    return 0;
}

/* MercurySolver ctor tail + Advance head: Sync(Badd), Sync(Bface),
 * UpdateDerivedFields_ (MercurySolver.cpp:141-145, MercurySolver_Driver.cpp:6-19) */
int orc_prepare(void *h) {
    Orc *o = (Orc *)h;
    sync_group(o, G_BADD);
    sync_group(o, G_BFACE);
    update_derived(o);
    return 0;
}

/* n steps; integrator 0 = forward Euler (reference StepOnce), 1 = SSP-RK2 */
int orc_step(void *h, int nsteps, int integrator, double *dt_out) {
    Orc *o = (Orc *)h;
    for (int n = 0; n < nsteps; ++n) {
        double dt = compute_timestep(o);
        if (!isfinite(dt) || dt < 1e-7) return -2;
        o->dt = dt;
        const int resist = o->P.resist_enabled != 0.0;
        int rc = 0;
        if (integrator == 0) {
            /* StepOnce (MercurySolver_Driver.cpp:21-61): the implicit resistive
             * correction follows the explicit update and Sync(Bface) */
            rc = stage(o, dt, resist);
        } else {
            /* SSP-RK2: the correction once per step, after the average */
            save_state(o);
            stage(o, dt, 0);
            stage(o, dt, 0);
            rk2_average(o);
            sync_group(o, G_UCELL);
            sync_group(o, G_BFACE);
            if (resist) rc = res_solve(o, dt, NULL, NULL);
            update_derived(o);
        }
        if (rc) return rc;
        o->time += dt;
        o->step += 1;
        if (dt_out) dt_out[n] = dt;
    }
    return 0;
}

/* DEC exterior derivatives on owned DOFs (src/8_dec/DecExteriorDerivative.cpp:
 * d0 node->edge, d1 edge->face = CurlEdgeToFace, d2 face->cell) */
static void dec_d0(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_EX + a, hi);
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k)
                        F(b, S_EDGEF + a, i, j, k) = F(b, S_NODEF, i + (a == 0), j + (a == 1), k + (a == 2)) - F(b, S_NODEF, i, j, k);
        }
    }
}
static void dec_d1(Orc *o, double mult) { curl_e2f(o, S_EDGEF, S_FACEF, mult); }
static void dec_d2(Orc *o) {
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k)
                    F(b, S_CELLS, i, j, k) = F(b, S_FACEF, i + 1, j, k) - F(b, S_FACEF, i, j, k) + F(b, S_FACEF + 1, i, j + 1, k) - F(b, S_FACEF + 1, i, j, k) +
                                             F(b, S_FACEF + 2, i, j, k + 1) - F(b, S_FACEF + 2, i, j, k);
    }
}

/* individual operators, for per-kernel parity tests */
int orc_op(void *h, const char *op, double arg) {
    Orc *o = (Orc *)h;
    if (!strcmp(op, "pv")) calc_pv(o);
    else if (!strcmp(op, "uplus")) calc_uplus(o);
    else if (!strcmp(op, "bcell")) calc_bcell(o);
    else if (!strcmp(op, "jedge")) calc_jedge(o);
    else if (!strcmp(op, "jcell")) calc_jcell(o);
    else if (!strcmp(op, "derived")) update_derived(o);
    else if (!strcmp(op, "flux")) scheme_u(o);
    else if (!strcmp(op, "sources")) add_sources(o);
    else if (!strcmp(op, "ideal_emf")) ideal_emf(o);
    else if (!strcmp(op, "hall_emf")) hall_emf(o);
    else if (!strcmp(op, "ambipolar_emf")) ambipolar_emf(o);
    else if (!strcmp(op, "arti_emf")) arti_emf(o);
    else if (!strcmp(op, "induction")) induction_ct(o);
    else if (!strcmp(op, "alias_e")) alias_reduce_edges(o, S_E);
    else if (!strcmp(op, "alias_j")) alias_reduce_edges(o, S_J);
    else if (!strcmp(op, "singular_e")) singular_emf(o, 0, S_E);
    else if (!strcmp(op, "singular_ehall")) singular_emf(o, 1, S_EH);
    else if (!strcmp(op, "singular_j")) singular_current(o, S_B, S_J);
    else if (!strcmp(op, "update")) apply_update(o, arg);
    else if (!strcmp(op, "sync_ucell")) sync_group(o, G_UCELL);
    else if (!strcmp(op, "sync_bface")) sync_group(o, G_BFACE);
    else if (!strcmp(op, "sync_eedge")) sync_group(o, G_EEDGE);
    else if (!strcmp(op, "sync_ehall")) sync_group(o, G_EHALL);
    else if (!strcmp(op, "sync_jedge")) sync_group(o, G_JEDGE);
    else if (!strcmp(op, "sync_bcell")) sync_group(o, G_BCELL);
    else if (!strcmp(op, "sync_jcell")) sync_group(o, G_JCELL);
    else if (!strcmp(op, "sync_badd")) sync_group(o, G_BADD);
    else if (!strcmp(op, "dec_d0")) dec_d0(o);
    else if (!strcmp(op, "dec_d1")) dec_d1(o, arg == 0.0 ? 1.0 : arg);
    else if (!strcmp(op, "dec_d2")) dec_d2(o);
    else if (!strcmp(op, "eface")) ideal_emf(o);
    else if (!strcmp(op, "res_matvec")) res_apply_fields(o, arg, 0);
    else if (!strcmp(op, "res_matvec_lin")) res_apply_fields(o, arg, 1);
    else if (!strcmp(op, "sync_eres")) sync_group(o, G_ERES);
    else if (!strcmp(op, "sync_db")) sync_group(o, G_DB);
    else if (!strcmp(op, "sync_dj")) sync_group(o, G_DJ);
    else return -1;
    return 0;
}

double orc_timestep(void *h) { return compute_timestep((Orc *)h); }

/* threads the per-point loops use */
int orc_threads(void) { return omp_get_max_threads(); }

/* one implicit resistive correction with time step dt on the current state;
 * returns 0, or -3 when the iteration limit is reached */
int orc_res_solve(void *h, double dt, int *its, double *rel) { return res_solve((Orc *)h, dt, its, rel); }

/* calc_divB (MercurySolver_MagneticCell.cpp:151-241) and the normalised
 * diagnostic max|d2 B| / max|B_face| over owned cells / faces */
double orc_divb(void *h, double *maxabs_out) {
    Orc *o = (Orc *)h;
    double dmaxv = 0.0, bmax = 0.0;
    for (int ib = 0; ib < o->nb; ++ib) {
        OBlock *b = &o->b[ib];
        for (int i = 0; i < b->m[0]; ++i)
            for (int j = 0; j < b->m[1]; ++j)
                for (int k = 0; k < b->m[2]; ++k) {
                    double d = (F(b, S_B, i + 1, j, k) - F(b, S_B, i, j, k) + F(b, S_B + 1, i, j + 1, k) - F(b, S_B + 1, i, j, k) +
                                F(b, S_B + 2, i, j, k + 1) - F(b, S_B + 2, i, j, k));
                    F(b, S_DIVB, i, j, k) = d / F(b, S_JAC, i, j, k);
                    dmaxv = dmax(dmaxv, fabs(d));
                }
        for (int a = 0; a < 3; ++a) {
            int hi[3];
            owned_hi(b, L_FX + a, hi);
            for (int i = 0; i < hi[0]; ++i)
                for (int j = 0; j < hi[1]; ++j)
                    for (int k = 0; k < hi[2]; ++k) bmax = dmax(bmax, fabs(F(b, S_B + a, i, j, k)));
        }
    }
    if (maxabs_out) *maxabs_out = dmaxv;
    return bmax > 0.0 ? dmaxv / bmax : dmaxv;
}

/* ------------------------------------------------------------------------- */
/* DEC operators of one block on caller arrays (src/8_dec)                     */
/* ------------------------------------------------------------------------- */
/* Array layout as above: N_a = m_a + 1 + 2g, k fastest.  Loops run over the
 * owned range of the output location, [0, m + 1 - delta) per axis
 * (FieldBlock::inner_lo / inner_hi, include/3_field/FieldBlock.h:34-35).
 * Pinned by the reference's own DEC chain test (d1 d0 = 0 and d2 d1 = 0 on its
 * polynomial fills, Z0_NULL/test/dec_chain/Z0_DecChainTest.cpp:13-51,85-100),
 * by single-point known answers (one unit value gives the signed incidence
 * pattern of the operator) and by max|.| of a single planted value. */
typedef struct {
    int m[3], g;
    long long sI, sJ;
} DecBox;
static DecBox dec_box(const int *dims, int g) {
    DecBox b;
    for (int a = 0; a < 3; ++a) b.m[a] = dims[a];
    b.g = g;
    b.sJ = dims[2] + 1 + 2LL * g;
    b.sI = (dims[1] + 1 + 2LL * g) * b.sJ;
    return b;
}
#define DQ(b, i, j, k) ((long long)((i) + (b).g) * (b).sI + (long long)((j) + (b).g) * (b).sJ + ((k) + (b).g))

/* DEC::d0_node_to_edge (src/8_dec/DecExteriorDerivative.cpp:41-77) */
void orc_dec_d0(const int *dims, int g, const double *n, double *ex, double *ey, double *ez) {
    const DecBox b = dec_box(dims, g);
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j <= b.m[1]; ++j)
            for (int k = 0; k <= b.m[2]; ++k) ex[DQ(b, i, j, k)] = n[DQ(b, i + 1, j, k)] - n[DQ(b, i, j, k)];
    for (int i = 0; i <= b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k <= b.m[2]; ++k) ey[DQ(b, i, j, k)] = n[DQ(b, i, j + 1, k)] - n[DQ(b, i, j, k)];
    for (int i = 0; i <= b.m[0]; ++i)
        for (int j = 0; j <= b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k) ez[DQ(b, i, j, k)] = n[DQ(b, i, j, k + 1)] - n[DQ(b, i, j, k)];
}

/* d1_xi_at / d1_eta_at / d1_zeta_at (DecExteriorDerivative.cpp:7-39) */
static double dec_d1_xi(DecBox b, const double *ey, const double *ez, int i, int j, int k) {
    return (ey[DQ(b, i, j, k)] - ey[DQ(b, i, j, k + 1)]) + (ez[DQ(b, i, j + 1, k)] - ez[DQ(b, i, j, k)]);
}
static double dec_d1_eta(DecBox b, const double *ex, const double *ez, int i, int j, int k) {
    return (ex[DQ(b, i, j, k + 1)] - ex[DQ(b, i, j, k)]) + (ez[DQ(b, i, j, k)] - ez[DQ(b, i + 1, j, k)]);
}
static double dec_d1_zeta(DecBox b, const double *ex, const double *ey, int i, int j, int k) {
    return (ex[DQ(b, i, j, k)] - ex[DQ(b, i, j + 1, k)]) + (ey[DQ(b, i + 1, j, k)] - ey[DQ(b, i, j, k)]);
}

/* DEC::d1_edge_to_face (:99-137); update != 0: faraday_update_face_2form,
 * face -= dt * d1(edge) (:202-241) */
void orc_dec_d1(const int *dims, int g, const double *ex, const double *ey, const double *ez, double *fx, double *fy, double *fz,
                int update, double dt) {
    const DecBox b = dec_box(dims, g);
    for (int i = 0; i <= b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k) {
                const double v = dec_d1_xi(b, ey, ez, i, j, k);
                if (update) fx[DQ(b, i, j, k)] -= dt * v;
                else fx[DQ(b, i, j, k)] = v;
            }
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j <= b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k) {
                const double v = dec_d1_eta(b, ex, ez, i, j, k);
                if (update) fy[DQ(b, i, j, k)] -= dt * v;
                else fy[DQ(b, i, j, k)] = v;
            }
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k <= b.m[2]; ++k) {
                const double v = dec_d1_zeta(b, ex, ey, i, j, k);
                if (update) fz[DQ(b, i, j, k)] -= dt * v;
                else fz[DQ(b, i, j, k)] = v;
            }
}

/* DEC::d2_face_to_cell (:164-180) */
void orc_dec_d2(const int *dims, int g, const double *fx, const double *fy, const double *fz, double *c) {
    const DecBox b = dec_box(dims, g);
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k)
                c[DQ(b, i, j, k)] = (fx[DQ(b, i + 1, j, k)] - fx[DQ(b, i, j, k)]) + (fy[DQ(b, i, j + 1, k)] - fy[DQ(b, i, j, k)]) +
                                    (fz[DQ(b, i, j, k + 1)] - fz[DQ(b, i, j, k)]);
}

/* DEC::max_abs_div_curl (src/8_dec/DecDiagnostics.cpp:110-135, curl_*_at :37-69):
 * sequential max from 0.0, std::max(acc, |x|) (a NaN value leaves acc unchanged) */
double orc_dec_max_abs_div_curl(const int *dims, int g, const double *ex, const double *ey, const double *ez) {
    const DecBox b = dec_box(dims, g);
    double acc = 0.0;
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k) {
                const double v = (dec_d1_xi(b, ey, ez, i + 1, j, k) - dec_d1_xi(b, ey, ez, i, j, k)) +
                                 (dec_d1_eta(b, ex, ez, i, j + 1, k) - dec_d1_eta(b, ex, ez, i, j, k)) +
                                 (dec_d1_zeta(b, ex, ey, i, j, k + 1) - dec_d1_zeta(b, ex, ey, i, j, k));
                acc = dmax(acc, fabs(v));
            }
    return acc;
}

/* DEC::max_abs_curl_grad (DecDiagnostics.cpp:71-93, abs_curl_grad_* :10-35) */
double orc_dec_max_abs_curl_grad(const int *dims, int g, const double *n) {
    const DecBox b = dec_box(dims, g);
#define NV(i, j, k) n[DQ(b, i, j, k)]
    double acc = 0.0;
    for (int i = 0; i <= b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k) {
                const double e0 = NV(i, j + 1, k) - NV(i, j, k), e1 = NV(i, j + 1, k + 1) - NV(i, j, k + 1);
                const double z0 = NV(i, j, k + 1) - NV(i, j, k), z1 = NV(i, j + 1, k + 1) - NV(i, j + 1, k);
                acc = dmax(acc, fabs((e0 - e1) + (z1 - z0)));
            }
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j <= b.m[1]; ++j)
            for (int k = 0; k < b.m[2]; ++k) {
                const double x0 = NV(i + 1, j, k) - NV(i, j, k), x1 = NV(i + 1, j, k + 1) - NV(i, j, k + 1);
                const double z0 = NV(i, j, k + 1) - NV(i, j, k), z1 = NV(i + 1, j, k + 1) - NV(i + 1, j, k);
                acc = dmax(acc, fabs((x1 - x0) + (z0 - z1)));
            }
    for (int i = 0; i < b.m[0]; ++i)
        for (int j = 0; j < b.m[1]; ++j)
            for (int k = 0; k <= b.m[2]; ++k) {
                const double x0 = NV(i + 1, j, k) - NV(i, j, k), x1 = NV(i + 1, j + 1, k) - NV(i, j + 1, k);
                const double e0 = NV(i, j + 1, k) - NV(i, j, k), e1 = NV(i + 1, j + 1, k) - NV(i + 1, j, k);
                acc = dmax(acc, fabs((x0 - x1) + (e1 - e0)));
            }
#undef NV
    return acc;
}
