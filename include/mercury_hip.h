/*
 * mercury_hip.h -- C ABI of the MI355X-native Mercury explicit step.
 *
 * The reference has no plugin/FFI layer: its "boundary" is the set of C++
 * class APIs that Z4_Mercury/main.cpp:27-99 wires together.  Each entry point
 * below replaces the reference interface named in its comment; all take plain
 * pointers and sizes, return 0 on success and -1 on failure (message in
 * mercury_last_error()), and never let an exception cross the boundary.
 *
 * Storage convention for every per-block array crossing this ABI: the box
 * [-g, m+g] on each axis of a block with m cells, k fastest,
 * element (i,j,k) at ((i+g)*(my+1+2g) + (j+g))*(mz+1+2g) + (k+g), one array per
 * field component.  It is the reference's FieldBlock(i,j,k,m) index space
 * (include/3_field/FieldBlock.h:38-49) with a common allocation for all
 * stagger locations.
 */
#ifndef MERCURY_HIP_H
#define MERCURY_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MERCURY_HIP_ABI_VERSION 1

/* Face boundary codes, order -xi,+xi,-eta,+eta,-zeta,+zeta.
 * PERIODIC: an inner self-connection with a negative source block id
 *           (src/1_grid/1_MPCNS_Grid.cpp:455-465).
 * FARFIELD: the "Farfield" physical patch (Z4_Mercury/src/boundary/Boundary_Build.cpp:45-203).
 * WALL:     the fluid side of the solid planet surface, i.e. the "Coupled-Solid"
 *           handlers (Boundary_Build.cpp:59-71, 111-119) and the stationary-wall EMF
 *           (Z4_Mercury/src/physics/MercurySolver_MagneticRHS.cpp:5-89).
 * INTERFACE: a 1:1 inner connection to a neighbour block (the reference's Inner
 *           InterfacePatch, include/2_topology/TopologyTypes.h:104-139), given to
 *           mercury_create_connected.
 * EXTERNAL: a face whose ghosts and shared DOFs the caller provides: no handler
 *           and no neighbour in the context; its halo copies come from
 *           mercury_set_sync_plan and the alias means of its shared edges from
 *           mercury_set_alias_classes (any connection the caller's own topology
 *           code can describe, rotated, mirrored or partial-face included). */
enum { MERCURY_BC_PERIODIC = 0, MERCURY_BC_FARFIELD = 1, MERCURY_BC_WALL = 2, MERCURY_BC_INTERFACE = 3, MERCURY_BC_EXTERNAL = 4 };

/* Integrators: forward Euler (the reference StepOnce) or SSP-RK2 (two Euler
 * stages and the average; no reference counterpart). */
enum { MERCURY_EULER = 0, MERCURY_SSPRK2 = 1 };

/* Non-dimensional constants, derived as in MercurySolver::calc_physical_constant
 * (Z4_Mercury/src/physics/MercurySolver_Tools.cpp:3-48) and
 * BuildMercuryBackgroundState (Z4_Mercury/include/0_BackgroundState.h:18-84). */
typedef struct MercuryPhysics {
    double gamma;
    double M_H, M_Na, M_ref;
    double state_coeff_H, state_coeff_Na;
    double T_ref;
    double inver_MA2;
    double hall_coef;
    double ne_floor;
    double c_ind, c_hall;            /* momentum_induce_coeff, momentum_hall_coeff */
    double a1_Na, a4, a5, a6;        /* AddSourceToRHS_Fluid coefficients */
    double Tn0;
    double nu_lit, nu_shadow;        /* drag frequencies sk1, sk2 */
    double cfl;
    double taper_rmin, taper_rmax;   /* Hall taper radii r_min, r_max */
    double arti_eta_max, J_start, J_on;
    double local_enabled, local_eta_max, local_center[3], local_r_decay, local_r_cutoff;
    double ambipolar_enabled, pressure_coef;
    double qinf[5], qinfs[5], q_pv_inf[5], B_imf[3];
    double wall_rho_floor_H, wall_rho_floor_Na, wall_p_floor;
    /* implicit resistive correction: is_Mercury_resistance, eta_max_mercury/(mu0 U_ref L_ref),
     * mercury_eta_r_inner/outer/width (MercurySolver.cpp:173-184, MercurySolver_Tools.cpp:44-69)
     * and the CG stopping rule |b - A x| <= max(rtol |b|, atol), at most max_it iterations */
    double resist_enabled, inver_Rem, resist_r_inner, resist_r_outer, resist_width;
    double resist_rtol, resist_atol, resist_max_it;
    /* eta profile: 0 the reference's window inver_Rem * 0.5 (tanh((r - r_inner) / width)
     * - tanh((r - r_outer) / width)) (MercurySolver_Tools.cpp:50-69); 1 an exponential
     * inver_Rem * exp(-(r - r_inner) / width) (no reference counterpart) */
    double resist_profile;
    /* ion species: 2 = H+ and Na+, the reference's two fluids
     * (Z4_Mercury/include/0_MercuryFieldCatalog.h:77-82); 3 adds He++ (no reference
     * counterpart) with its molar mass, state coefficient R T_ref / (M U_ref^2),
     * charge number (its charge density Z n joins n_e, the U+ weights, the Hall and
     * induction forces), far-field state and wall density floor.  Its state is the
     * field U_He (PV_He, RHS_He); sums over species add the He++ term last. */
    double n_species, M_He, state_coeff_He, Z_He, qinf_He[5], wall_rho_floor_He;
} MercuryPhysics;

typedef struct MercuryContext MercuryContext;

const char *mercury_last_error(void);
int mercury_version(void);

/* Replaces Grid::Grid_Preprocess (include/1_grid/1_MPCNS_Grid.h:152-226) +
 * Field(Grid*,Param*,ngg) with METRIC::build_field_geometry (src/7_metric/Metric.cpp:1416)
 * + MercurySolver::RegisterFields/RegisterHaloFields and the solver constructor's
 * static setup (Z4_Mercury/src/MercurySolver.cpp:11-247, SetupCellReconstructionWeights_).
 * dims: 3 ints per block (cells); bc: 6 codes per block; owned_nodes: per block
 * x, y, z planes of the (mx+1)(my+1)(mz+1) owned nodes, k fastest. */
int mercury_create(MercuryContext **ctx, int nblocks, const int *dims, int nghost, const int *bc,
                   const double *owned_nodes, const MercuryPhysics *phys);

/* mercury_create for blocks glued by 1:1 interfaces: the topology that
 * TOPO::build_topology (src/2_topology/TopologyBuilder.cpp:436) derives from the
 * grid's inner-connection files.  For every face with code MERCURY_BC_INTERFACE,
 * nbr[6*b+f] is the neighbour block and transform[9*(6*b+f) ...] its
 * IndexTransform (perm[3], sign[3], offset[3]: neighbour node index[perm[a]] =
 * sign[a] * node index[a] + offset[a], include/2_topology/TopologyTypes.h:93-98);
 * other faces ignore both.  Supported are translations (perm identity, sign +1)
 * that glue a face to the opposite face of the neighbour over its whole extent,
 * declared from both sides, with at least nghost cells along the connection;
 * anything else fails with a message.  A block may be its own neighbour (a
 * periodic axis).  Halo copies, owner -> alias copies and the edge alias means
 * then run across blocks as in the reference's staged Sync_
 * (Z4_Mercury/src/boundary/Boundary_Sync_.cpp:4-110).  nbr and transform may be
 * NULL when no face is an interface. */
int mercury_create_connected(MercuryContext **ctx, int nblocks, const int *dims, int nghost, const int *bc, const int *nbr,
                             const int *transform, const double *owned_nodes, const MercuryPhysics *phys);

/* mercury_create for blocks joined by patches: any axis permutation and
 * reversal, faces split into several patches, a block joined to itself for
 * periodicity; the topology TOPO::build_topology derives from the grid's
 * inner-connection files (src/2_topology/TopologyBuilder.cpp:64-260, patches
 * read by mercury_case_interfaces).  patches: npatch records of 18 ints
 * (block, direction +-1..3, node box lo[3] hi[3] on the block face (hi
 * exclusive), neighbour block, perm[3], sign[3], offset[3]: neighbour node
 * index[perm[a]] = sign[a] * node index[a] + offset[a]), in the order of the
 * reference's inner patch list (it decides which face fills a corner).  Faces
 * with code MERCURY_BC_INTERFACE must be covered by their patches; other codes
 * are MERCURY_BC_FARFIELD or MERCURY_BC_WALL (no MERCURY_BC_PERIODIC: periodicity
 * is a patch).  The halo replays the reference's three levels of regions
 * (Halo::build_inner_{1,2,3}DCorner_pattern, src/4_halo/build/HaloBuildFace.cpp,
 * HaloBuildEdge.cpp, HaloBuildVertex.cpp; copies of HaloExchangeInner.cpp and
 * src/4_halo/forms/HaloTriplet.cpp) and runs as one flat copy list per group
 * and stage (mercury_sync_plan reads them back); shared faces and edges take
 * their owner's value with the topology's orientation signs
 * (src/2_topology/Topology.cpp:280-323).
 * rule 0 reproduces the reference: staggered DOFs are mapped as node indices,
 * face 2-forms take the factor sign[0]*sign[1]*sign[2]*sign[src] and read, at
 * the edge and vertex levels, the axis a with perm[a] = the destination axis.
 * Along a reversed axis such a DOF lands one index away from the DOF at its
 * position, an odd permutation flips the 2-form factor, and corner ghosts the
 * reference's regions skip keep older values.  rule 1 maps every DOF by
 * position, carries the area vector's orientation (permutation parity) and
 * fills the ghosts no region writes through chains of patches: on the
 * reference's sample geometries a step then equals the one-block cube's. */
int mercury_create_patched(MercuryContext **ctx, int nblocks, const int *dims, int nghost, const int *bc, int npatch,
                           const int *patches, int rule, const double *owned_nodes, const MercuryPhysics *phys);

/* Host only (no GPU, no context): the copies mercury_create_patched derives
 * for one kind of field group (0 cell component copies, 1 face 2-form triplet,
 * 2 edge 1-form triplet) up to a halo level (1 face strips, 3 with edges and
 * vertices), stage 0 the halo and stage 1 owner -> alias.  Entries of 11 ints:
 * destination (field axis, block, i, j, k), source (field axis, block, i, j, k),
 * sign; a halo source is the DOF's value when the halo starts, so an entry list
 * whose sources include its destinations is applied with every read before
 * every write.  *n receives the count; entries are written when capacity >= *n. */
int mercury_patch_plan(int nblocks, const int *dims, int nghost, const int *bc, int npatch, const int *patches, int rule, int kind,
                       int level, int stage, long long capacity, long long *n, int *entries);
/* Host only (no GPU, no context): the singular edges mercury_create_patched
 * finds (SingularEdgeRegistry::build, src/7_metric/SingularEdgeRegistry.cpp:
 * 78-268): shared edges whose aliases touch three, or more than four, owner
 * faces, as on the corner lines of a cubed sphere.  On them the context
 * assembles E and E_hall from the cells around every alias
 * (Z4_Mercury/src/physics/MercurySolver_MagExplicit_Ideal.cpp:57-130,
 * MercurySolver_MagExplicit_Hall.cpp:32-56) and J and dJ from the owner faces
 * around them (MercurySolver_Current.cpp:59-138), and alpha on their aliases is
 * length / dual area, the dual area being the polygon through the cells'
 * centres.  Per entry 8 ints: owner (block, i, j, k, axis), the owner alias's
 * orientation, cell count, face count; and 4 doubles: the inverse Hodge scale
 * and the mean edge vector.  *n receives the count; entries are written when
 * capacity >= *n.  Fails as the reference aborts for an entry without length,
 * dual area or cell. */
int mercury_singular_edges(int nblocks, const int *dims, int nghost, const int *bc, int npatch, const int *patches, int rule,
                           const double *owned_nodes, const MercuryPhysics *phys, int capacity, int *n, int *ints,
                           double *values);
/* An ensemble of independent runs in ONE context (no reference counterpart: the
 * reference runs one case per MPI job): run r owns the next run_nblocks[r] blocks
 * of the concatenated block list (dims, bc, owned_nodes as for
 * mercury_create_connected; nbr holds indices into the concatenated list and an
 * interface may not join two runs) and phys[r] is its physics.  The runs may
 * differ only in hall_coef, the far-field states (qinf, qinfs, q_pv_inf,
 * qinf_He) and B_imf; everything else must be equal, and the implicit resistive
 * correction is refused.  Every pass of a step is one launch over every block
 * of every run; each run has its own CFL time step and its own dt abort status
 * (a refused run stays at its last completed step, mercury_step and
 * mercury_graph_launch fail naming the run).  Fields are addressed by the
 * block's index in the concatenated list. */
int mercury_create_ensemble(MercuryContext **ctx, int nruns, const int *run_nblocks, const int *dims, int nghost, const int *bc,
                            const int *nbr, const int *transform, const double *owned_nodes, const MercuryPhysics *phys);
/* Number of runs of a context (1 unless made by mercury_create_ensemble) and,
 * when dts is not NULL, the dt each run committed for its last step (nruns
 * doubles; 0 for a run stopped by the abort rule).  mercury_step's dt_out and
 * mercury_timestep report run 0. */
int mercury_run_dt(MercuryContext *ctx, int *nruns, double *dts);
/* Replaces the deletes at the end of Z4_Mercury/main.cpp:86-89. */
int mercury_destroy(MercuryContext *ctx);

/* Replaces Field::field_id / descriptor (include/3_field/Field.h:16-96). */
int mercury_field_info(const char *name, int *slot, int *ncomp, int *location);
/* Elements of one per-block array (the box volume). */
long long mercury_block_volume(MercuryContext *ctx, int block);
long long mercury_total_volume(MercuryContext *ctx);

/* Cells per axis and ghost layers of one block (FieldBlock::get_block / the
 * Block extents, include/3_field/FieldBlock.h:27-35). */
int mercury_block_shape(MercuryContext *ctx, int block, int *dims, int *nghost);
/* Device address of component comp of a field on one block: the block box of
 * the storage convention above, mercury_block_volume(ctx, block) doubles,
 * owned by the context and valid until mercury_destroy.  It is the device
 * counterpart of FieldBlock::raw_data (include/3_field/FieldBlock.h:52) and
 * lets a caller run its own kernels or the mercury_dec_* operators on solver
 * fields in place.  Waits for the context's stream first; in the mixed form a
 * field the step holds in FP32 is first widened into its FP64 array. */
int mercury_field_device(MercuryContext *ctx, const char *name, int block, int comp, double **ptr);
/* The context's HIP stream (a hipStream_t): work a caller orders against the
 * step goes on it. */
int mercury_stream(MercuryContext *ctx, void **stream);
/* Halo plans as flat int32 lists (no single reference counterpart: they carry
 * what Halo's patterns do, include/4_halo/HaloTypes.h:8-35, applied as
 * Halo::exchange_inner / the triplet copies, src/4_halo/component_copy/
 * HaloExchangeInner.cpp:90-247, src/4_halo/forms/HaloTriplet.cpp:641-725, and
 * the owner -> alias copies, src/4_halo/core/HaloSyncFramework.cpp:695-860).
 * group: a sync group of MercuryBoundary::Sync ("Ucell", "Bface", "Eedge",
 * "Ehall", "Jedge", "B_cell", "J_cell", "Badd", "Eres", "Eres0" (Eres with a
 * zero far field), "dB").  stage 0: the halo copies, stage 1: the owner ->
 * alias copies, which run after them.  An entry is 11 int32:
 *   dst slot, dst block, di, dj, dk, src slot, src block, si, sj, sk, sign
 * with slot = the slot of mercury_field_info + the component, points (i, j, k)
 * in the reference's index space of the slot's location (ghosts negative or
 * beyond the owned range, inside the block's box), sign +1 or -1; the entry
 * sets dst = sign * src, and the entries of one stage act as if every source
 * were read before any destination is written (a stage whose sources include
 * its own destinations runs as a gather and a scatter through a scratch list).
 * Setting a stage replaces the context's own copies of that stage for that
 * group (the other stage keeps what it had); the physical handlers and corner
 * defaults of the group still run first.  Drops a captured step graph.
 * mercury_sync_plan writes the stage's current entries (the context's own
 * until a caller sets them) and their number; with entries NULL or capacity
 * below the number, only the number. */
int mercury_set_sync_plan(MercuryContext *ctx, const char *group, int stage, long long n, const int *entries);
int mercury_sync_plan(MercuryContext *ctx, const char *group, int stage, long long capacity, long long *n, int *entries);
/* The classes of ReduceEdgeAliasCandidatesToOwners_ (Z4_Mercury/src/physics/
 * MercurySolver_Current.cpp:3-57): for class c, edges of axis axis[c] (0 xi,
 * 1 eta, 2 zeta), members start[c] .. start[c+1]-1 of members (5 int32 each:
 * block, i, j, k, sign), owner owners[5c .. 5c+4]; the E and J edge passes set
 * owner = owner sign * (sum of member sign * member) / count.  Setting replaces
 * the context's own classes; mercury_alias_classes writes the current ones
 * (counts only when an array is NULL or too small).  Drops a captured step
 * graph. */
int mercury_set_alias_classes(MercuryContext *ctx, int ncls, const int *start, const int *axis, const int *members, const int *owners);
int mercury_alias_classes(MercuryContext *ctx, int cls_capacity, long long member_capacity, int *ncls, long long *nmembers, int *start,
                          int *axis, int *members, int *owners);

/* Run the context's work on a caller's HIP stream (a hipStream_t; NULL returns
 * to the context's own stream).  No reference counterpart: the reference's
 * solver is synchronous host code; this is the stream of the drop-in's context
 * (steps, graph launches, syncs, uploads/downloads, DEC diagnostics).  Work
 * already queued on the previous stream completes before the call returns.
 * The caller keeps the stream alive while the context uses it and gives it the
 * default priority; the fork of the fused step still uses an internal side
 * stream, joined back by events. */
int mercury_set_stream(MercuryContext *ctx, void *stream);

/* Replace direct FieldBlock(i,j,k,m) access on host storage
 * (include/3_field/FieldBlock.h:38-49) and the restart injection of
 * IOModule::ReadRestartBinFile (src/6_io/IOModule_Bin.cpp:234-420). */
int mercury_upload(MercuryContext *ctx, const char *name, int block, int comp, const double *host);
int mercury_download(MercuryContext *ctx, const char *name, int block, int comp, double *host);

/* Initial face fluxes: B = Bu . JD + d1(A) and Badd = B_imf . JD on the allocated
 * face range; A holds 3 edge arrays per block (or NULL).  Badd is the IMF part of
 * DipoleField::Build_Badd_FaceFlux (Z4_Mercury/include/operators/Dipole.h:326-372). */
int mercury_init_bface(MercuryContext *ctx, const double *Bu, const double *const *A);
/* mercury_init_bface plus the intrinsic dipoles of DipoleField::Build_Badd_FaceFlux
 * (Dipole.h:89-374, called from Mercury_Initial::Initial_Badd, Z4_Mercury/include/2_Initial.h:219-224):
 * Badd += d1(A_dip . dr) on the face boxes of ghost depth 1, where A_dip is the
 * sum over dipoles of (m x r) / (|r|^2 + a^2)^(3/2), r = x - r0, taken at each
 * edge midpoint (DipoleField::A_of_dipole, :25-39).  dipoles holds 7
 * nondimensional doubles per dipole, x0 y0 z0 mx my mz a, as
 * DipoleField::load_from_param (:44-87) derives them from the Dipole<i> lists
 * (r0 and a over L_ref, m over 4 pi B_ref L_ref^3 / mu_mag).  Needs nghost >= 1. */
int mercury_init_bface_dipoles(MercuryContext *ctx, const double *Bu, const double *const *A, int ndipoles,
                               const double *dipoles);

/* Replaces the constructor tail Sync("Badd"), Sync("Bface") (MercurySolver.cpp:141-145)
 * and the UpdateDerivedFields_ call at the head of Advance (MercurySolver_Driver.cpp:6-19).
 * Also clears a stop of the dt abort rule (below). */
int mercury_prepare(MercuryContext *ctx);

/* Replaces the while(!if_stop) StepOnce() loop of MercurySolver::Advance
 * (Z4_Mercury/src/MercurySolver_Driver.cpp:6-61) for nsteps steps; dt_out (nsteps
 * doubles, may be NULL) receives each step's dt.
 * The reference's dt abort rule (MercurySolver_Timestep.cpp:129-140: dt not
 * finite or below 1e-7 ends the run before the update) is applied on the device
 * at the start of every step, in mercury_step, mercury_graph_launch and
 * mercury_ensemble_launch alike: the refused step and every later one leave the
 * state untouched, and the call returns -1 with a message naming the refused
 * step (counted since mercury_prepare).  The context refuses further steps until
 * mercury_prepare is called again. */
int mercury_step(MercuryContext *ctx, int nsteps, int integrator, double *dt_out);

/* Individual operators (one reference function each), for parity tests and
 * drop-in use: pv_uplus, bcell, jedge, jcell, derived, flux, sources, eface,
 * edge_emf, hall_cell, hall_edge, induction, alias_e, alias_j, timestep, stage,
 * dec_d0, dec_d1 (arg = multiplier), dec_d2, sync_ucell, sync_bface, sync_eedge,
 * sync_ehall, sync_jedge, sync_bcell, sync_jcell, sync_badd, divb (calc_divB's
 * reduction left on the device; mercury_divb reads it). */
int mercury_op(MercuryContext *ctx, const char *op, double arg);

/* Replaces MercurySolver::Compute_Timestep (MercurySolver_Timestep.cpp:8-141). */
int mercury_timestep(MercuryContext *ctx, double *dt);

/* Replaces calc_divB (MercurySolver_MagneticCell.cpp:151-241); returns
 * max|d2 B| / max|B_face| and max|d2 B| over owned cells/faces. */
int mercury_divb(MercuryContext *ctx, double *normalised, double *maxabs);
/* The per-step div B monitor (no reference counterpart; it reports calc_divB's
 * figure without a pass of its own): the fused forms' B-cell pass, which reads
 * the synchronised face fluxes of every owned cell anyway, folds max|d2 B| and
 * max|B_face| into per-workgroup maxima, and one small kernel turns them into
 * the normalised value of mercury_divb.  last: that value for the latest
 * derived pass (the end of the latest step), max_since_reset: its maximum over
 * every derived pass since the last reset (Euler: every step; SSP-RK2: every
 * stage), last_maxabs: the latest max|d2 B|.  reset != 0 clears them after the
 * read.  Any pointer may be NULL.  The plain form (0) does not update it.
 * mercury_set_divb_monitor switches it on (off by default: on shell6 it adds
 * about 12 us to the B-cell pass); switching drops a captured step graph. */
int mercury_set_divb_monitor(MercuryContext *ctx, int on);
int mercury_divb_monitor(MercuryContext *ctx, double *last, double *max_since_reset, double *last_maxabs, int reset);

int mercury_synchronize(MercuryContext *ctx);

/* Form of the step: 0 = one kernel per reference function (the plain FP64
 * drop-in), 1 = fused FP64 passes (default; identical results to 0),
 * 2 = mixed precision: fluxes, EMFs, Hall terms and reconstructions in FP32
 * (compensated flux divergence) with FP32 static arrays; state, metric,
 * CT update, floors and dt in FP64. */
int mercury_set_form(MercuryContext *ctx, int form);

/* Implicit resistive correction (is_Mercury_resistance; replaces
 * MercurySolver::SolveImplicitResistiveDiffusion_, Mercury_ImplicitResistive.cpp:467-547,
 * with ApplyImplicitResistiveUpdate_ :549-575).  mercury_step runs it after
 * every Euler update (once per SSP-RK2 step) when the physics struct enables
 * it; op "res_solve" (arg = dt) runs one correction on the current state.
 * The reference's PETSc GMRES becomes conjugate gradients in the
 * diag(eta*alpha)^-1 inner product with a Jacobi preconditioner.  Reports the
 * number of unknown edges and, for the last solve, the iterations and the
 * relative residual |b - A x| / |b|. */
int mercury_resist_stats(MercuryContext *ctx, int *n_unknowns, int *iterations, double *rel_residual);

/* One explicit step captured as a HIP graph and replayed (no reference
 * counterpart: the reference loops on the host).  Not available with the
 * implicit resistive correction (its iteration count is data dependent).
 * mercury_graph_launch replays nsteps times, waits, and applies the dt abort
 * rule as mercury_step does.  mercury_set_form and mercury_autotune drop the
 * captured graph. */
int mercury_graph_capture(MercuryContext *ctx, int integrator);
int mercury_graph_launch(MercuryContext *ctx, int nsteps);

/* Restart of device-resident fields in the reference's MPCNSRST version-1 file
 * (IOModule::WriteRestartBinFile, src/6_io/IOModule_Bin.cpp:168-228): the
 * reference's whitelist U_H, U_Na, B_xi, B_eta, B_zeta (Z4_Mercury/src/MercurySolver.cpp:31-44),
 * plus J_xi, J_eta, J_zeta when with_jedge is nonzero (its OUTPUT_DEC_JEDGE
 * option); every block active, values over the allocated box of the location in
 * (i, j, k, component) order. */
int mercury_write_restart(MercuryContext *ctx, const char *path, int step, double time, int with_jedge);

/* IOModule::ReadRestartBinFile (src/6_io/IOModule_Bin.cpp:230-420): magic,
 * version and block count are checked; every field the library knows by name is
 * validated (location, ncomp, nghost, per-block lo/hi) and uploaded, others are
 * stepped over; a partial J edge triplet or an inactive block of a known field is
 * an error.  step and time return the file's values.  Call mercury_prepare
 * afterwards (the reference syncs Badd and Bface and rebuilds the derived fields
 * after a restart, Z4_Mercury/src/MercurySolver.cpp:135-145). */
int mercury_read_restart(MercuryContext *ctx, const char *path, int *step, double *time);

/* RunData::WriteBinary / ReadBinary (include/6_io/RunData.h:98-205), the
 * MPCNSRUN version-1 file.  Monitors are written empty; on reading they are
 * parsed and dropped.  mercury_read_rundata stores at most max_ref residual
 * references (an error if the file holds more) and returns their count in nref. */
int mercury_write_rundata(const char *path, int step, double time, double dt, int nref, const double *residual_ref,
                          double residual_max);
int mercury_read_rundata(const char *path, int *step, double *time, double *dt, int max_ref, int *nref, double *residual_ref,
                         double *residual_max);

/* IOModule::WriteTecplotBinFile in the all-node mode the Mercury solver selects
 * (src/6_io/IOModule_Tecplot.cpp:691-1017; Z4_Mercury/src/MercurySolver.cpp:53-86):
 * a Tecplot binary (#!TDV112) file titled "flow_field" with one ordered zone
 * "Zone<b>_<name>" per block over its (m0+1) x (m1+1) x (m2+1) owned nodes,
 * variables x, y, z and the components of the listed cell or node fields, all
 * written at the nodes as FP32 with FP64 min/max.  A cell value reaches a node
 * as the |Jac|-weighted mean of the 8 cells around it, ghosts included
 * (EvalValue_CellToNode_, :322-418), computed on the device.  fields = NULL
 * (nfields = 0) writes the Mercury list PV_H, PV_Na, B_cell, Bind_cell, Na,
 * J_cell with its component names (u_H ... Jz); unknown names and fields at
 * faces or edges are skipped with a message, as in the reference.
 * block_names = NULL names every zone's block "Fluid". */
int mercury_write_tecplot(MercuryContext *ctx, const char *path, double time, const char *const *fields, int nfields,
                          const char *const *block_names);

/* CASE geometry directories: Grid::ReadGeometry (src/1_grid/1_MPCNS_Grid.cpp:10-590)
 * over the files of every rank ("grid   r.grd", boundary_condition/"boundary   r.txt",
 * boundary_name.fvbnd, "inner   r.txt", "parallel   r.txt"; r padded to 4
 * characters), merged into one block list in rank-major order, and the patches
 * of TOPO::build_topology (src/2_topology/TopologyBuilder.cpp:436-462): inner
 * patches as written, parallel patches paired by their send/receive flags in
 * the target rank's file, each with the IndexTransform of
 * make_index_transform_from_boundary_arrays (src/2_topology/TopologyOps.cpp:118-219).
 * nrank <= 0 counts the grid files present; binary selects Read_Grid_Binary
 * (Read_Method != 0) or the text Read_Grid. */
typedef struct MercuryCase MercuryCase;
int mercury_case_read(MercuryCase **c, const char *geometry_dir, int nrank, int binary);
void mercury_case_destroy(MercuryCase *c);
/* block count and, when dims is not NULL, cells per axis (nodes - 1) per block */
int mercury_case_blocks(const MercuryCase *c, int *nblocks, int *dims);
int mercury_case_block_name(const MercuryCase *c, int block, char *buf, int len);
/* owned node coordinates in the layout mercury_create takes */
int mercury_case_nodes(const MercuryCase *c, double *owned_nodes);
/* bc / nbr / transform for mercury_create_connected: every face must be one
 * patch over the whole face, a Farfield or Solid_Surface boundary (codes
 * MERCURY_BC_FARFIELD, MERCURY_BC_WALL) or a connection (MERCURY_BC_INTERFACE,
 * periodic ones included); anything else fails with a message naming the face. */
int mercury_case_topology(const MercuryCase *c, int *bc, int *nbr, int *transform);
/* bc and patches for mercury_create_patched (any connection patches: split,
 * rotated, mirrored, periodic): faces holding connection patches get
 * MERCURY_BC_INTERFACE, the others must be one whole-face Farfield or
 * Solid_Surface patch; *npatch receives the number of 18-int patch records,
 * written to patches when capacity >= *npatch (the reader's order). */
int mercury_case_patch_topology(const MercuryCase *c, int *bc, int capacity, int *npatch, int *patches);
/* patch listings (records may be NULL to query n).  Connection patch, 27 ints:
 * kind (0 inner, 1 parallel), periodic, block, nb_block, direction,
 * nb_direction (+-(axis+1)), node box lo[3] hi[3], neighbour node box lo[3]
 * hi[3] (half-open), perm[3], sign[3], offset[3].  Physical patch, 9 ints:
 * block, direction, bc id (boundary_name.fvbnd), node box lo[3] hi[3]. */
int mercury_case_interfaces(const MercuryCase *c, int *n, int *records);
int mercury_case_physical(const MercuryCase *c, int *n, int *records);

/* Ensemble of independent runs (no reference counterpart: the reference runs
 * one case per MPI job).  One HIP graph forks into every context's stream,
 * captures one step of each (integrator as in mercury_graph_capture) and joins
 * back; mercury_ensemble_launch replays it nsteps times, waits, and reports the
 * first member stopped by the dt abort rule.  Contexts keep their own physics,
 * dt and state and must outlive the ensemble; contexts with the implicit
 * resistive correction are refused.  A member whose form or launch shape
 * changed after the capture (mercury_set_form, mercury_autotune) makes
 * mercury_ensemble_launch fail before anything is launched. */
typedef struct MercuryEnsemble MercuryEnsemble;
int mercury_ensemble_create(MercuryContext *const *ctxs, int n, int integrator, MercuryEnsemble **out);
int mercury_ensemble_launch(MercuryEnsemble *ens, int nsteps);
int mercury_ensemble_destroy(MercuryEnsemble *ens);

/* Launch-shape autotuner (no reference counterpart): times the flux pass
 * (Scheme_U_ + AddSourceToRHS_Fluid) in the current form as the column kernel
 * with 1, 2, 4 and 8 cells per thread along i and as the LDS-tiled kernel with
 * 4, 8 and 16 planes per 16 x 16 tile, and keeps the fastest; results are unchanged by the choice.  Call after
 * mercury_set_form; it drops a captured step graph (capture again afterwards).
 * Writes the chosen length (query the kernel with mercury_flux_shape). */
int mercury_autotune(MercuryContext *ctx, int *march_out);

/* The flux pass's launch shape (no reference counterpart): kind 0 runs the
 * column kernel with length cells per thread along i, kind 1 the LDS-tiled
 * kernel with length planes per tile.  Both give the same result bit for bit.  Setting it
 * drops a captured step graph; length < 1 or another kind is refused. */
int mercury_set_flux_shape(MercuryContext *ctx, int kind, int length);
int mercury_flux_shape(MercuryContext *ctx, int *kind, int *length);

/* Average device time (ms) of an op over reps launches, HIP events on the
 * solver's stream; op "step_graph" replays the captured step. */
int mercury_time_op(MercuryContext *ctx, const char *op, int reps, double *ms_avg);

/* Host-only static geometry of one block (no GPU needed): writes component
 * comp of the named metric / reconstruction field (JDxi, Jac, beta_*, alpha_*,
 * dr_*, Bcell_from_Bface_w, ...) over the block box.  Replaces reading the
 * fields of METRIC::compute_metric_fields (src/7_metric/Metric.cpp:78-1414) and
 * SetupCellReconstructionWeights_ (Z4_Mercury/src/MercurySolver_ReconstructionSetup.cpp:29-193). */
int mercury_geometry_block(const int *dims, int nghost, const double *owned_nodes, const MercuryPhysics *phys,
                           const char *name, int comp, double *out);

/* Discrete exterior derivatives on one block held in CALLER device arrays
 * (no context): dims = cells per axis, nghost = ghost layers, every array one
 * component over the block box of the storage convention above.  Each
 * operator writes the owned range of its output location ([0, m + 1 - delta)
 * per axis, FieldBlock::inner_lo/inner_hi, include/3_field/FieldBlock.h:34-35)
 * and reads its inputs there and one index above; ghost layers are left
 * untouched.  stream is a hipStream_t (NULL: the null stream); the calls are
 * asynchronous on it.  edge and face are arrays of 3 pointers (xi, eta, zeta). */
/* DEC::d0_node_to_edge(FieldBlock&...) (include/8_dec/DecOps.h:11-16,
 * src/8_dec/DecExteriorDerivative.cpp:41-77) */
int mercury_dec_d0(const int *dims, int nghost, const double *node, double *const *edge, void *stream);
/* DEC::d1_edge_to_face(FieldBlock&...) (DecOps.h:24-31, DecExteriorDerivative.cpp:99-137) */
int mercury_dec_d1(const int *dims, int nghost, const double *const *edge, double *const *face, void *stream);
/* DEC::d2_face_to_cell(FieldBlock&...) (DecOps.h:39-44, DecExteriorDerivative.cpp:164-180) */
int mercury_dec_d2(const int *dims, int nghost, const double *const *face, double *cell, void *stream);
/* DEC::faraday_update_face_2form(FieldBlock&...): face -= dt * d1(edge)
 * (DecOps.h:52-60, DecExteriorDerivative.cpp:202-241) */
int mercury_dec_faraday_update(const int *dims, int nghost, const double *const *edge, double *const *face, double dt,
                               void *stream);
/* DEC::max_abs_div_curl(FieldBlock&...) (include/8_dec/DecDiagnostics.h:17-20,
 * src/8_dec/DecDiagnostics.cpp:110-135): max over owned cells of |d2 d1 edge|,
 * written to result, ONE DEVICE double.  The per-workgroup partial maxima live
 * in a stream-ordered allocation (hipMallocAsync / hipFreeAsync on stream). */
int mercury_dec_max_abs_div_curl(const int *dims, int nghost, const double *const *edge, double *result, void *stream);
/* DEC::max_abs_curl_grad(FieldBlock&...) (DecDiagnostics.h:11, DecDiagnostics.cpp:71-93):
 * max over owned faces of |d1 d0 node|, to one device double, as above. */
int mercury_dec_max_abs_curl_grad(const int *dims, int nghost, const double *node, double *result, void *stream);
/* mercury_dec_max_abs_div_curl without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The partial maxima go to a static device array of the library (1 MiB per
 * GPU, held until the process ends), so a block may have at most 33 554 432
 * owned cells (one partial per 256 cells); larger blocks fail with a message.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
int mercury_dec_max_abs_div_curl_noalloc(const int *dims, int nghost, const double *const *edge, double *result,
                                         void *stream);

#ifdef __cplusplus
}
#endif
#endif
