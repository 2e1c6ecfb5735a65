// Shared host/device definitions of the MI355X Mercury explicit step.
//
// Storage model (device): every field component is one contiguous FP64 array
// over the concatenation of all blocks.  Block b occupies [off_b, off_b + V_b)
// where V_b = N0*N1*N2, N_a = m_a + 1 + 2g, and point (i,j,k), i,j,k in
// [-g, m+g], sits at off_b + ((i+g)*N1 + (j+g))*N2 + (k+g) (k fastest, SoA).
// All stagger locations share the box, so a (block, i, j, k) -> index map and
// every halo/boundary index list is valid for every field of a location.
// This mirrors the reference's FieldBlock(i,j,k,m) accessor
// (include/3_field/FieldBlock.h:38-49) with the component hoisted out of the
// innermost position (AoS -> SoA) for coalesced wave64 access along k.
#pragma once
#include <stdint.h>

#define MERC_MAX_SYNC_SLOTS 10

// stagger locations (include/0_basic/StaggerLocation.h)
enum MercLoc { LOC_CELL = 0, LOC_FX, LOC_FY, LOC_FZ, LOC_EX, LOC_EY, LOC_EZ, LOC_NODE };

// field slots ---------------------------------------------------------------
enum MercSlot {
    // cell
    SL_UH = 0, SL_UN = SL_UH + 5, SL_PVH = SL_UN + 5, SL_PVN = SL_PVH + 5,
    SL_UP = SL_PVN + 5, SL_BC = SL_UP + 3, SL_BI = SL_BC + 3, SL_JC = SL_BI + 3,
    SL_RH = SL_JC + 3, SL_RN = SL_RH + 5, SL_EHC = SL_RN + 5, SL_HB = SL_EHC + 3,
    SL_JAC = SL_HB + 1, SL_WB = SL_JAC + 1, SL_WJ = SL_WB + 18, SL_PHOTO = SL_WJ + 36,
    SL_CC = SL_PHOTO + 1, SL_DIVB = SL_CC + 3, SL_U0H = SL_DIVB + 1, SL_U0N = SL_U0H + 5,
    SL_CELLS = SL_U0N + 5,
    // faces
    SL_B = SL_CELLS + 1, SL_BADD = SL_B + 3, SL_RHSB = SL_BADD + 3, SL_JD = SL_RHSB + 3,
    SL_AREA = SL_JD + 9, SL_BETA = SL_AREA + 3, SL_EF = SL_BETA + 3, SL_F = SL_EF + 9,
    SL_B0 = SL_F + 15, SL_FACEF = SL_B0 + 3,
    // edges
    SL_E = SL_FACEF + 3, SL_EH = SL_E + 3, SL_J = SL_EH + 3, SL_DR = SL_J + 3, SL_DL = SL_DR + 9,
    SL_ALPHA = SL_DL + 3, SL_TAPER = SL_ALPHA + 3, SL_LETA = SL_TAPER + 3, SL_EDGEF = SL_LETA + 3,
    // nodes
    SL_NODEF = SL_EDGEF + 3,
    // cell: contravariant velocity of U+ per axis (fused path)
    SL_UCON = SL_NODEF + 1,
    // implicit resistive correction: Eres, dJ and eta on the unknown edges;
    // dB and RHS_Bres faces
    SL_ERES = SL_UCON + 3, SL_DJ = SL_ERES + 3, SL_ETAR = SL_DJ + 3, SL_DB = SL_ETAR + 3, SL_RHSBR = SL_DB + 3,
    // node: electron pressure (mean of p_H + p_Na over the 8 cells around the
    // node) of the ambipolar EMF (fused path)
    SL_PEN = SL_RHSBR + 3,
    // cell: neutral Na density, a static input that the step does not read
    // (the reference's "Na" field, written by the output)
    SL_NANEU = SL_PEN + 1,
    // the third ion species (He++): state, PV, RHS and step-start copy
    SL_UHE = SL_NANEU + 1, SL_PVHE = SL_UHE + 5, SL_RHE = SL_PVHE + 5, SL_U0HE = SL_RHE + 5,
    SL_NSLOT = SL_U0HE + 5
};

// species order H+, Na+, He++ (the reference's two first: every sum over
// species adds the He++ term last, so two-species runs are unchanged)
#define MERC_NSMAX 3
#define SPEC_U(s) ((s) == 0 ? SL_UH : ((s) == 1 ? SL_UN : SL_UHE))
#define SPEC_PV(s) ((s) == 0 ? SL_PVH : ((s) == 1 ? SL_PVN : SL_PVHE))
#define SPEC_R(s) ((s) == 0 ? SL_RH : ((s) == 1 ? SL_RN : SL_RHE))
#define SPEC_U0(s) ((s) == 0 ? SL_U0H : ((s) == 1 ? SL_U0N : SL_U0HE))

// physics constants (same layout as MercuryPhysics in include/mercury_hip.h)
struct MercPhys {
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
    // implicit resistive correction (MercurySolver.cpp:173-184, MercurySolver_Tools.cpp:44-69)
    double resist_enabled, inver_Rem, resist_r_inner, resist_r_outer, resist_width;
    double resist_rtol, resist_atol, resist_max_it;
    double resist_profile;  // 0: the reference's tanh window, 1: exponential
    // ion species: 2 (H+, Na+) or 3 with He++ (no reference counterpart)
    double n_species, M_He, state_coeff_He, Z_He, qinf_He[5], wall_rho_floor_He;
};

// per-block geometry of the box
struct MercBlk {
    long long off;   // offset of the block in every field array
    int m[3];        // cells per axis
    int g;           // ghost depth
    int N[3];        // box extent m+1+2g
    int sI, sJ;      // strides of i and j (k stride is 1)
    int bc[6];       // face codes, -xi,+xi,-eta,+eta,-zeta,+zeta
    int run;         // run of the block in an ensemble context (0 otherwise): its dt and status index
    double hall_coef;  // the run's Hall coefficient (the one physics constant the kernels read per run)
};

// a launch range: per block a box [lo, lo+n) of (i,j,k), linearised k fastest;
// start[b] is the first thread of block b (start[nb] = total)
struct MercRange {
    int nb;
    const long long *start;  // device, nb+1
    const int *lo;           // device, 3*nb
    const int *n;            // device, 3*nb
};

// face codes; MBC_INTERFACE: 1:1 connection to a neighbour block (node-index translation)
// MBC_EXTERNAL: a face whose ghosts (and shared DOFs) come from a caller-supplied
// sync plan and alias classes; the context adds no handler and no neighbour there
enum MercBC { MBC_PERIODIC = 0, MBC_FARFIELD = 1, MBC_WALL = 2, MBC_INTERFACE = 3, MBC_EXTERNAL = 4 };

// slot sets of a merged sync op (passed by value as a kernel argument)
#define MERC_MAX_SETS 12
struct MultiSets {
    int n[MERC_MAX_SETS];
    int s[MERC_MAX_SETS][MERC_MAX_SYNC_SLOTS];
    double v[MERC_MAX_SETS][MERC_MAX_SYNC_SLOTS];
};
